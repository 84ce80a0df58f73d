// C-ABI of the hot path: argument validation, device memory, dispatch.
// Every argument is validated on the host before any HIP call; errors are
// returned as vrs_status with a message (the reference prints and exit(1)s,
// error.cuh:4-18, or prints and continues, vissGen10M_shade_optim2.cu:208-211).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "host_sky.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "io.hpp"
#include "refcompat.hpp"
#include "grid.hpp"
#include "vrs.h"
#include "vrs_internal.hpp"

using namespace vrs;

namespace {
thread_local std::string g_thread_err;
constexpr double kC = 299792458.0;
}  // namespace

// Scratch of vrs_recon_noalloc: the partial images of its (map, visibility
// segment) slots, and in its last kNoallocQueueBytes the recon kernel's work
// counter.  Each device that loads the library holds its own copy at its
// own address until the process ends.  Host code never uses this name as an
// address: it takes the device address once per device with
// hipGetSymbolAddress (noalloc_scratch_addr) and passes that pointer on.
static __device__ __align__(256) float scratch[VRS_NOALLOC_SCRATCH_BYTES / sizeof(float)];

namespace {
constexpr size_t kNoallocQueueBytes = 256;
std::mutex g_scratch_mu;
std::map<int, float *> g_scratch_addr;  // device -> address of its copy of `scratch`

// The device address of `scratch` on the current device (`device`), looked up
// once per device.
hipError_t noalloc_scratch_addr(int device, float **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = (float *)p;
    *out = (float *)p;
    return hipSuccess;
}
}  // namespace

struct vrs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int n_cu = 0;
    std::string err;
    bool timing = false;
    std::map<std::string, std::pair<hipEvent_t, hipEvent_t>> events;
    // grow-only device workspaces
    void *ws[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t ws_bytes[4] = {0, 0, 0, 0};
    unsigned long long *d_tot = nullptr;
    // autotuned batch tiles, keyed by (padded visibilities, pixels, channels, precision)
    std::map<std::tuple<int64_t, int64_t, int32_t, int32_t>, std::pair<int, int>> tuned;
    // autotuned single-geometry vis-gen tiles, keyed by (n_vis, n_pix, n_maps, precision)
    std::map<std::tuple<int64_t, int64_t, int32_t, int32_t>, int> tuned_vt;
    // device address of this device's copy of the allocation-free recon's scratch
    float *noalloc_scratch = nullptr;
};

struct vrs_geom {
    vrs_ctx *ctx = nullptr;  // owning context; must outlive the geometry
    int device = 0;
    int32_t n_sat = 0, n_t = 0, n_base = 0, nside = 0, n_chunk = 0;
    int32_t n_half = 0;  // > 0: chunk k + n_half holds the antipodes of chunk k, slot by slot
    int64_t n_vis = 0, n_pix = 0, n_pad = 0;
    double lambda_m = 0, r_moon = 0, freq_hz = 0;
    int32_t uvw_flags = 0;
    bool from_uvw = false;  // built from an explicit uvw row list: no satellites, no orbits
    OrbitArgs orb{};
    std::vector<int32_t> pairs;
    std::vector<double> lmn_ring;  // [n_pix][3] host copy, RING order
    // device
    double *d_orbit = nullptr;  // raan[n_sat], m0[n_sat]
    int64_t *d_tidx = nullptr;  // [n_t] sample index of each step, or null
    double4 *d_pos = nullptr;
    CapRec *d_cap = nullptr;
    double4 *d_uvw64 = nullptr;
    int2 *d_viscap = nullptr;
    int2 *d_pairs = nullptr;
    float4 *d_pix32 = nullptr;
    double4 *d_pix64 = nullptr;
    int32_t *d_ring = nullptr;
    ChunkRec *d_chunk = nullptr;
    ChunkRec *d_sub = nullptr;  // bounding caps of the kSubChunk-pixel groups
    // fine-cell sky (vrs_geom_fine_sky); all null until one is attached
    int64_t n_cells = 0;
    float4 *d_cell32 = nullptr;
    double4 *d_cell64 = nullptr;
    int32_t *d_fidx = nullptr, *d_fstart = nullptr, *d_fcount = nullptr;
    float *d_cbar = nullptr;
    float2 *d_pixcap = nullptr;
    ChunkRec *d_fchunk = nullptr;
    void free_fine() {
        void *ptrs[] = {d_cell32, d_cell64, d_fidx, d_fstart, d_fcount, d_cbar, d_pixcap, d_fchunk};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        d_cell32 = nullptr; d_cell64 = nullptr; d_fidx = d_fstart = d_fcount = nullptr;
        d_cbar = nullptr; d_pixcap = nullptr; d_fchunk = nullptr;
        n_cells = 0;
    }
    GeomDev dev() const {
        GeomDev g = {};
        g.n_sat = n_sat; g.n_t = n_t; g.n_base = n_base; g.n_chunk = n_chunk; g.n_half = n_half;
        g.n_vis = n_vis; g.n_pad = n_pad;
        g.pos = d_pos; g.cap = d_cap; g.uvw64 = d_uvw64; g.vis_cap = d_viscap; g.pairs = d_pairs;
        g.pix32 = d_pix32; g.pix64 = d_pix64; g.ring = d_ring; g.chunk = d_chunk;
        g.sub = d_sub;
        const double e = kNearLimbRel * r_moon;
        g.limb.R2 = r_moon * r_moon;
        g.limb.lo2 = (r_moon - e) * (r_moon - e);
        g.limb.hi2 = (r_moon + e) * (r_moon + e);
        return g;
    }
};

namespace {

vrs_status fail(vrs_ctx *ctx, vrs_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    g_thread_err = buf;
    return st;
}

#define VRS_HIP(ctx, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((ctx), e_ == hipErrorOutOfMemory ? VRS_E_OOM : VRS_E_HIP,           \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,    \
                        __LINE__);                                                          \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

bool is_device_ptr(const void *p) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

vrs_status ws_get(vrs_ctx *ctx, int slot, size_t bytes, void **out) {
    if (ctx->ws_bytes[slot] < bytes) {
        if (ctx->ws[slot]) {
            VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
            VRS_HIP(ctx, hipFree(ctx->ws[slot]));
            ctx->ws[slot] = nullptr;
            ctx->ws_bytes[slot] = 0;
        }
        size_t b = std::max(bytes, (size_t)256);
        VRS_HIP(ctx, hipMalloc(&ctx->ws[slot], b));
        ctx->ws_bytes[slot] = b;
    }
    *out = ctx->ws[slot];
    return VRS_OK;
}

// A caller array that may live on the host: used in place when it is device
// memory, otherwise staged through a private device buffer.
struct Staged {
    vrs_ctx *ctx;
    const void *user;
    void *dev = nullptr;
    size_t bytes;
    bool host = false;
    bool out;
    Staged(vrs_ctx *c, const void *u, size_t b, bool is_out) : ctx(c), user(u), bytes(b), out(is_out) {}
    vrs_status prepare() {
        if (!user) return VRS_OK;
        if (is_device_ptr(user)) {
            dev = const_cast<void *>(user);
            return VRS_OK;
        }
        host = true;
        VRS_HIP(ctx, hipMalloc(&dev, std::max(bytes, (size_t)4)));
        if (!out) VRS_HIP(ctx, hipMemcpyAsync(dev, user, bytes, hipMemcpyHostToDevice, ctx->stream));
        return VRS_OK;
    }
    vrs_status finish() {
        if (!host || !dev) return VRS_OK;
        if (out)
            VRS_HIP(ctx, hipMemcpyAsync(const_cast<void *>(user), dev, bytes,
                                        hipMemcpyDeviceToHost, ctx->stream));
        return VRS_OK;
    }
    ~Staged() {
        if (host && dev) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(dev);
        }
    }
};

struct Timer {
    vrs_ctx *ctx;
    std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
    Timer(vrs_ctx *c, const char *name) : ctx(c) {
        if (!ctx->timing) return;
        auto it = ctx->events.find(name);
        if (it == ctx->events.end()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess) return;
            if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
            it = ctx->events.emplace(name, std::make_pair(a, b)).first;
        }
        ev = &it->second;
        (void)hipEventRecord(ev->first, ctx->stream);
    }
    void stop() {
        if (ev) (void)hipEventRecord(ev->second, ctx->stream);
        ev = nullptr;
    }
    ~Timer() { stop(); }
};

int persistent_grid(vrs_ctx *ctx, const void *kernel, int64_t n_items, int block = kBlock) {
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    int64_t cap = (int64_t)per_cu * ctx->n_cu;
    int64_t grid = std::min<int64_t>(n_items, cap);
    if (grid >= 8) grid -= grid % 8;  // whole rounds over the 8 XCDs
    return (int)std::max<int64_t>(grid, 1);
}

vrs_status check_ctx(vrs_ctx *ctx) {
    if (!ctx) return fail(nullptr, VRS_E_ARG, "ctx is NULL");
    return VRS_OK;
}

vrs_status check_knobs(vrs_ctx *ctx, const vrs_knobs *k) {
    if (!k) return fail(ctx, VRS_E_ARG, "knobs is NULL");
    if (!std::isfinite(k->phase_scale) || k->phase_scale == 0.0)
        return fail(ctx, VRS_E_ARG, "phase_scale must be finite and non-zero (got %g)", k->phase_scale);
    if (k->wfac != 0 && k->wfac != 1) return fail(ctx, VRS_E_ARG, "wfac must be 0 or 1");
    if (!std::isfinite(k->t_occ)) return fail(ctx, VRS_E_ARG, "t_occ must be finite");
    if (!std::isfinite(k->kappa) || k->kappa < 0.0)
        return fail(ctx, VRS_E_ARG, "kappa must be finite and >= 0 (0 selects 1/N_vis)");
    if (k->precision < VRS_PLAIN_F32 || k->precision > VRS_MFMA_SPLIT16)
        return fail(ctx, VRS_E_ARG, "unknown precision %d", k->precision);
    if (k->antipodal != 0 && k->antipodal != 1) return fail(ctx, VRS_E_ARG, "antipodal must be 0 or 1");
    if (k->sat_factors != 0 && k->sat_factors != 1)
        return fail(ctx, VRS_E_ARG, "sat_factors must be 0 or 1");
    return VRS_OK;
}

// Autotuner cache key part: precision and antipodal pairing.
int32_t tune_mode(const vrs_knobs *k) { return k->precision + 16 * k->antipodal; }

PhaseKnobs phase_knobs(const vrs_knobs *k) {
    PhaseKnobs p;
    p.s = k->phase_scale;
    p.wfac = k->wfac;
    p.t_occ = k->t_occ;
    p.pair = k->antipodal;
    return p;
}

}  // namespace

extern "C" {

const char *vrs_version(void) { return "vrs 0.1.0 (gfx950)"; }

void vrs_knobs_default(vrs_knobs *k) {
    if (!k) return;
    k->phase_scale = 1.0;
    k->wfac = 0;
    k->t_occ = 240.0f;
    k->kappa = 0.0;
    k->precision = VRS_PLAIN_F32;
    k->antipodal = 1;
    k->sat_factors = 1;
}

void vrs_knobs_reference(vrs_knobs *k) {
    if (!k) return;
    k->phase_scale = 0.5;  // dl = 2*RES/(RES-1) == 2 in integer arithmetic
    k->wfac = 1;
    k->t_occ = 240.0f;
    k->kappa = 0.0;
    k->precision = VRS_PLAIN_F32;
    k->antipodal = 1;
    k->sat_factors = 1;
}

int vrs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

const char *vrs_last_error(const vrs_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_thread_err.c_str();
}

vrs_status vrs_create(int device, vrs_ctx **out) {
    if (!out) return fail(nullptr, VRS_E_ARG, "out is NULL");
    *out = nullptr;
    int n = vrs_device_count();
    if (n <= 0) return fail(nullptr, VRS_E_NODEV, "no HIP device available");
    if (device < 0 || device >= n)
        return fail(nullptr, VRS_E_ARG, "device %d out of range [0, %d)", device, n);
    DeviceGuard dg(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return fail(nullptr, VRS_E_HIP, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, VRS_E_NODEV, "device %d is %s; this library is built for gfx950",
                    device, prop.gcnArchName);
    vrs_ctx *c = new vrs_ctx;
    c->device = device;
    c->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, VRS_E_HIP, "hipStreamCreate failed");
    }
    c->own_stream = true;
    // [0, 4): shade statistics; [4]: the recon work counter (context stream)
    if (hipMalloc(&c->d_tot, 8 * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, VRS_E_OOM, "hipMalloc of counters failed");
    }
    if (hipError_t e = noalloc_scratch_addr(device, &c->noalloc_scratch)) {
        (void)hipFree(c->d_tot);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, VRS_E_HIP, "hipGetSymbolAddress of the recon scratch failed: %s",
                    hipGetErrorString(e));
    }
    *out = c;
    return VRS_OK;
}

void vrs_destroy(vrs_ctx *ctx) {
    if (!ctx) return;
    DeviceGuard dg(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->events) {
        (void)hipEventDestroy(kv.second.first);
        (void)hipEventDestroy(kv.second.second);
    }
    for (int i = 0; i < 4; ++i)
        if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
    if (ctx->d_tot) (void)hipFree(ctx->d_tot);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

vrs_status vrs_set_stream(vrs_ctx *ctx, void *stream) {
    if (vrs_status s = check_ctx(ctx)) return s;
    DeviceGuard dg(ctx->device);
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) {
        VRS_HIP(ctx, hipStreamDestroy(ctx->stream));
        ctx->own_stream = false;
    }
    // NULL selects the device's null stream (the legacy default stream)
    ctx->stream = (hipStream_t)stream;
    return VRS_OK;
}

vrs_status vrs_synchronize(vrs_ctx *ctx) {
    if (vrs_status s = check_ctx(ctx)) return s;
    DeviceGuard dg(ctx->device);
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VRS_OK;
}

vrs_status vrs_timing_enable(vrs_ctx *ctx, int on) {
    if (vrs_status s = check_ctx(ctx)) return s;
    ctx->timing = on != 0;
    return VRS_OK;
}

double vrs_timing_last_ms(vrs_ctx *ctx, const char *kernel) {
    if (!ctx || !kernel) return -1.0;
    auto it = ctx->events.find(kernel);
    if (it == ctx->events.end()) return -1.0;
    DeviceGuard dg(ctx->device);
    if (hipEventSynchronize(it->second.second) != hipSuccess) return -1.0;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, it->second.first, it->second.second) != hipSuccess) return -1.0;
    return ms;
}

vrs_status vrs_term_ceiling(vrs_ctx *ctx, double *gterms_per_s) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!gterms_per_s) return fail(ctx, VRS_E_ARG, "gterms_per_s is required");
    DeviceGuard dg(ctx->device);
    // eight blocks of 256 per CU; about 2 ms per launch on an MI355X
    const int32_t blocks = 8 * std::max(ctx->n_cu, 1), iters = 2048;
    struct Buf {
        void *p = nullptr;
        ~Buf() { if (p) (void)hipFree(p); }
    } b_uvw, b_out;
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() { if (e) (void)hipEventDestroy(e); }
    } e0, e1;
    float4 tab[16];
    for (int i = 0; i < 16; ++i)  // (u, v, w, w offset) in revolutions, |phase| up to ~1e5
        tab[i] = make_float4(1.0e5f - 977.0f * i, -3.1e4f + 411.0f * i, 812.5f + 3.0f * i,
                             -812.5f - 3.0f * i);
    VRS_HIP(ctx, hipMalloc(&b_uvw.p, sizeof(tab)));
    VRS_HIP(ctx, hipMalloc(&b_out.p, sizeof(float) * 256 * (size_t)blocks));
    VRS_HIP(ctx, hipMemcpyAsync(b_uvw.p, tab, sizeof(tab), hipMemcpyHostToDevice, ctx->stream));
    VRS_HIP(ctx, hipEventCreate(&e0.e));
    VRS_HIP(ctx, hipEventCreate(&e1.e));
    double best = 0.0;
    for (int rep = 0; rep < 3; ++rep) {  // the first launch warms the clocks
        VRS_HIP(ctx, hipEventRecord(e0.e, ctx->stream));
        VRS_HIP(ctx, launch_term_ceiling((const float4 *)b_uvw.p, iters, blocks, (float *)b_out.p,
                                         ctx->stream));
        VRS_HIP(ctx, hipEventRecord(e1.e, ctx->stream));
        VRS_HIP(ctx, hipEventSynchronize(e1.e));
        float ms = 0.0f;
        VRS_HIP(ctx, hipEventElapsedTime(&ms, e0.e, e1.e));
        if (rep > 0 && ms > 0.0f)
            best = std::max(best, (double)term_ceiling_terms(iters, blocks) / (ms * 1e-3) / 1e9);
    }
    *gterms_per_s = best;
    return VRS_OK;
}

uint32_t vrs_pix_hash(int64_t p) { return pix_hash_u64((uint64_t)p); }

// ---------------------------------------------------------------- host helpers
vrs_status vrs_rng_normals(uint64_t seed, int64_t n, double *out) {
    if (n < 0 || (n > 0 && !out)) return fail(nullptr, VRS_E_ARG, "bad n or out");
    rng_normals(seed, n, out);
    return VRS_OK;
}

vrs_status vrs_rng_uniforms(uint64_t seed, int64_t n, double *out) {
    if (n < 0 || (n > 0 && !out)) return fail(nullptr, VRS_E_ARG, "bad n or out");
    rng_uniforms(seed, n, out);
    return VRS_OK;
}

vrs_status vrs_sky_lognormal(uint64_t seed, double sigma, int64_t n_pix, int32_t n_src,
                             double s_min, double s_max, float *out) {
    if (n_pix <= 0 || !out) return fail(nullptr, VRS_E_ARG, "n_pix must be > 0 and out non-NULL");
    if (!std::isfinite(sigma) || sigma < 0) return fail(nullptr, VRS_E_ARG, "sigma must be >= 0");
    if (n_src < 0 || n_src > n_pix) return fail(nullptr, VRS_E_ARG, "n_src out of range");
    if (n_src > 0 && !(s_min > 0 && s_max > s_min))
        return fail(nullptr, VRS_E_ARG, "need 0 < s_min < s_max");
    sky_lognormal(seed, sigma, n_pix, n_src, s_min, s_max, out);
    return VRS_OK;
}

vrs_status vrs_orbit_offsets(int32_t n_sat, uint64_t seed, double a_km, double max_offset_km,
                             double *raan, double *m0) {
    if (n_sat <= 0 || !raan || !m0) return fail(nullptr, VRS_E_ARG, "bad n_sat or outputs");
    if (!(a_km > 0) || !(max_offset_km >= 0) || max_offset_km > 2 * a_km)
        return fail(nullptr, VRS_E_ARG, "need a_km > 0 and 0 <= max_offset_km <= 2 a_km");
    orbit_offsets(n_sat, seed, a_km, max_offset_km, raan, m0);
    return VRS_OK;
}

vrs_status vrs_uvw_position_index(int32_t orbit_res, int32_t orbit_counts, int32_t interval,
                                  int32_t span, int64_t *out) {
    if (orbit_res < 3 || orbit_counts < 1 || interval < 0 || span < 1 || !out)
        return fail(nullptr, VRS_E_ARG, "need orbit_res >= 3, orbit_counts >= 1, interval >= 0, "
                                        "span >= 1 and out");
    if (span != orbit_res / 3)
        return fail(nullptr, VRS_E_ARG, "span must be orbit_res / 3 (each output orbit block is "
                                        "orbit_res / 3 steps long)");
    if ((int64_t)orbit_res * orbit_counts > ((int64_t)1 << 40))
        return fail(nullptr, VRS_E_ARG, "orbit_res * orbit_counts too large");
    uvw_position_index(orbit_res, orbit_counts, interval, span, out);
    return VRS_OK;
}

vrs_status vrs_healpix_lmn(int32_t nside, double *lmn) {
    if (nside < 1 || nside > 8192 || !lmn) return fail(nullptr, VRS_E_ARG, "bad nside or lmn");
    const int64_t n = 12LL * nside * nside;
    for (int64_t p = 0; p < n; ++p) healpix_ring_vec(nside, p, lmn + 3 * p);
    return VRS_OK;
}

vrs_status vrs_healpix_nest2ring(int32_t nside, int64_t *out) {
    if (nside < 1 || nside > 8192 || (nside & (nside - 1)) || !out)
        return fail(nullptr, VRS_E_ARG, "nside must be a power of two in [1, 8192]");
    const int64_t n = 12LL * nside * nside;
    for (int64_t p = 0; p < n; ++p) out[p] = healpix_nest2ring(nside, p);
    return VRS_OK;
}

vrs_status vrs_healpix_antipode(int32_t nside, int64_t *out) {
    if (nside < 1 || nside > 8192 || (nside & (nside - 1)) || !out)
        return fail(nullptr, VRS_E_ARG, "nside must be a power of two in [1, 8192]");
    const int64_t n = 12LL * nside * nside;
    for (int64_t p = 0; p < n; ++p) out[p] = healpix_ring_antipode(nside, p);
    return VRS_OK;
}

// ---------------------------------------------------------------- geometry
void vrs_geom_destroy(vrs_geom *g) {
    if (!g) return;
    DeviceGuard dg(g->device);
    (void)hipDeviceSynchronize();
    void *ptrs[] = {g->d_orbit, g->d_tidx, g->d_pos,  g->d_cap,  g->d_uvw64, g->d_viscap,
                    g->d_pairs, g->d_pix32, g->d_pix64, g->d_ring, g->d_chunk, g->d_sub};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    g->free_fine();
    delete g;
}

// The sky of a geometry: pixel directions in chunk order (HEALPix NESTED for a
// HEALPix sky, so that a chunk is a compact patch; the given order for explicit
// directions), their RING / row index, and each chunk's bounding cap.
static vrs_status sky_build(vrs_ctx *ctx, vrs_geom *g, int32_t nside, int64_t n_pix,
                            const double *lmn) {
    g->nside = lmn ? 0 : nside;
    g->n_pix = lmn ? n_pix : 12LL * nside * nside;
    g->n_chunk = (int32_t)((g->n_pix + kChunk - 1) / kChunk);
    g->n_pad = (int64_t)g->n_chunk * kChunk;
    std::vector<float4> pix32(g->n_pad, make_float4(0, 0, 0, 0));
    std::vector<double4> pix64(g->n_pad, make_double4(0, 0, 0, 0));
    std::vector<int32_t> ring(g->n_pad, -1);
    g->lmn_ring.resize(3 * g->n_pix);
    if (lmn)
        std::memcpy(g->lmn_ring.data(), lmn, sizeof(double) * 3 * g->n_pix);
    else
        for (int64_t p = 0; p < g->n_pix; ++p) healpix_ring_vec(nside, p, &g->lmn_ring[3 * p]);
    // Chunk order: NESTED (16 x 16 blocks of a base face are chunks).  From
    // Nside 16 on, the second half is re-ordered so that position n/2 + q holds
    // the antipode of position q: base faces 0-5 face faces 6-11, so chunk
    // k + n_chunk/2 is the antipodal image of chunk k, slot by slot.
    g->n_half = 0;
    const bool paired = !lmn && nside >= 16;
    std::vector<int64_t> order(g->n_pix);
    for (int64_t q = 0; q < g->n_pix; ++q) order[q] = lmn ? q : healpix_nest2ring(nside, q);
    if (paired) {
        const int64_t h = g->n_pix / 2;
        std::vector<uint8_t> used(g->n_pix, 0);
        bool ok = (g->n_pix % (2 * kChunk)) == 0;
        for (int64_t q = 0; q < h && ok; ++q) {
            const int64_t a = healpix_ring_antipode(nside, order[q]);
            order[h + q] = a;
        }
        for (int64_t q = 0; q < g->n_pix && ok; ++q) {
            if (used[order[q]]) ok = false;
            used[order[q]] = 1;
        }
        if (ok) {
            g->n_half = g->n_chunk / 2;
        } else {
            for (int64_t q = 0; q < g->n_pix; ++q) order[q] = healpix_nest2ring(nside, q);
        }
    }
    for (int64_t q = 0; q < g->n_pix; ++q) {
        const int64_t r = order[q];
        const double *v = &g->lmn_ring[3 * r];
        pix64[q] = make_double4(v[0], v[1], v[2], 0.0);
        pix32[q] = make_float4((float)v[0], (float)v[1], (float)v[2], 0.0f);
        ring[q] = (int32_t)r;
    }
    // Bounding cap of the pixels [q0, q0 + len): every pixel centre lies within
    // angle rho of (gx, gy, gz); rho is padded, and wide or partly padded groups
    // are flagged to be tested pixel by pixel.
    auto bound = [&](int64_t q0, int len) {
        ChunkRec ch;
        double sx = 0, sy = 0, sz = 0;
        int nv = 0;
        uint32_t h = 0;
        for (int i = 0; i < len; ++i) {
            const int64_t q = q0 + i;
            if (ring[q] < 0) continue;
            sx += pix64[q].x; sy += pix64[q].y; sz += pix64[q].z;
            ++nv;
            h ^= pix_hash_u64((uint64_t)ring[q]);
        }
        const double nrm = std::sqrt(sx * sx + sy * sy + sz * sz);
        double rho = M_PI;
        if (nrm > 1e-6 * nv) {
            sx /= nrm; sy /= nrm; sz /= nrm;
            double maxchord = 0;
            for (int i = 0; i < len; ++i) {
                const int64_t q = q0 + i;
                if (ring[q] < 0) continue;
                const double dx = pix64[q].x - sx, dy = pix64[q].y - sy, dz = pix64[q].z - sz;
                maxchord = std::max(maxchord, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
            rho = 2.0 * std::asin(std::min(1.0, 0.5 * maxchord));
            rho = rho * (1.0 + 1e-6) + 1e-6;
        } else {
            sx = 0; sy = 0; sz = 1;
        }
        ch.gx = (float)sx; ch.gy = (float)sy; ch.gz = (float)sz;
        ch.crho = (float)std::cos(rho);
        ch.srho = (float)std::sin(rho);
        ch.flags = (rho > 0.5 || nv < len) ? kChunkAlwaysStraddle : 0;
        ch.n_valid = nv;
        ch.hash_all = h;
        return ch;
    };
    std::vector<ChunkRec> chunks(g->n_chunk), subs(g->n_pad / kSubChunk);
    for (int32_t c = 0; c < g->n_chunk; ++c) chunks[c] = bound((int64_t)c * kChunk, kChunk);
    for (size_t i = 0; i < subs.size(); ++i) subs[i] = bound((int64_t)i * kSubChunk, kSubChunk);

    hipStream_t st = ctx->stream;
    VRS_HIP(ctx, hipMalloc(&g->d_pix32, sizeof(float4) * g->n_pad));
    VRS_HIP(ctx, hipMalloc(&g->d_pix64, sizeof(double4) * g->n_pad));
    VRS_HIP(ctx, hipMalloc(&g->d_ring, sizeof(int32_t) * g->n_pad));
    VRS_HIP(ctx, hipMalloc(&g->d_chunk, sizeof(ChunkRec) * g->n_chunk));
    VRS_HIP(ctx, hipMalloc(&g->d_sub, sizeof(ChunkRec) * subs.size()));
    VRS_HIP(ctx, hipMemcpyAsync(g->d_sub, subs.data(), sizeof(ChunkRec) * subs.size(),
                                hipMemcpyHostToDevice, st));
    VRS_HIP(ctx, hipMemcpyAsync(g->d_pix32, pix32.data(), sizeof(float4) * g->n_pad,
                                hipMemcpyHostToDevice, st));
    VRS_HIP(ctx, hipMemcpyAsync(g->d_pix64, pix64.data(), sizeof(double4) * g->n_pad,
                                hipMemcpyHostToDevice, st));
    VRS_HIP(ctx, hipMemcpyAsync(g->d_ring, ring.data(), sizeof(int32_t) * g->n_pad,
                                hipMemcpyHostToDevice, st));
    VRS_HIP(ctx, hipMemcpyAsync(g->d_chunk, chunks.data(), sizeof(ChunkRec) * g->n_chunk,
                                hipMemcpyHostToDevice, st));
    // the host-side copies above are pageable; finish them before the vectors go away
    VRS_HIP(ctx, hipStreamSynchronize(st));
    return VRS_OK;
}

static vrs_status geom_build(vrs_ctx *ctx, const vrs_geom_desc *d, vrs_geom *g) {
    const vrs_orbit_desc &o = d->orbit;
    g->n_sat = o.n_sat;
    g->n_t = o.n_t;
    g->r_moon = d->r_moon_km;
    g->uvw_flags = d->uvw_flags;
    g->lambda_m = (d->uvw_flags & VRS_UVW_LAMBDA_INT) ? (double)(int)(3e8 / d->freq_hz)
                                                      : kC / d->freq_hz;
    g->freq_hz = d->freq_hz;
    if (d->n_base == 0) {
        for (int a = 0; a < o.n_sat; ++a)
            for (int b = a + 1; b < o.n_sat; ++b) {
                g->pairs.push_back(a);
                g->pairs.push_back(b);
            }
    } else {
        g->pairs.assign(d->pairs, d->pairs + 2 * d->n_base);
    }
    g->n_base = (int32_t)(g->pairs.size() / 2);
    g->n_vis = (int64_t)g->n_base * g->n_t;
    if (vrs_status s = sky_build(ctx, g, d->nside, 0, nullptr)) return s;

    hipStream_t st = ctx->stream;
    const int64_t nst = (int64_t)g->n_sat * g->n_t;
    VRS_HIP(ctx, hipMalloc(&g->d_orbit, 2 * sizeof(double) * g->n_sat));
    VRS_HIP(ctx, hipMalloc(&g->d_pos, sizeof(double4) * nst));
    VRS_HIP(ctx, hipMalloc(&g->d_cap, sizeof(CapRec) * nst));
    VRS_HIP(ctx, hipMalloc(&g->d_uvw64, sizeof(double4) * std::max<int64_t>(g->n_vis, 1)));
    VRS_HIP(ctx, hipMalloc(&g->d_viscap, sizeof(int2) * std::max<int64_t>(g->n_vis, 1)));
    VRS_HIP(ctx, hipMalloc(&g->d_pairs, sizeof(int2) * g->n_base));
    if (o.sample_index) {
        VRS_HIP(ctx, hipMalloc(&g->d_tidx, sizeof(int64_t) * g->n_t));
        VRS_HIP(ctx, hipMemcpyAsync(g->d_tidx, o.sample_index, sizeof(int64_t) * g->n_t,
                                    hipMemcpyHostToDevice, st));
    }

    std::vector<double> orbit(2 * g->n_sat);
    for (int k = 0; k < g->n_sat; ++k) {
        orbit[k] = o.raan_rad[k];
        orbit[g->n_sat + k] = o.m0_rad[k];
    }
    VRS_HIP(ctx, hipMemcpyAsync(g->d_orbit, orbit.data(), sizeof(double) * orbit.size(),
                                hipMemcpyHostToDevice, st));
    VRS_HIP(ctx, hipMemcpyAsync(g->d_pairs, g->pairs.data(), sizeof(int32_t) * g->pairs.size(),
                                hipMemcpyHostToDevice, st));

    OrbitArgs &a = g->orb;
    a.n_sat = g->n_sat;
    a.n_t = g->n_t;
    a.dt = o.dt_s;
    a.a = o.a_km;
    a.cos_i = std::cos(o.incl_rad);
    a.sin_i = std::sin(o.incl_rad);
    a.n_mot = std::sqrt(o.mu_km3_s2 / (o.a_km * o.a_km * o.a_km));
    a.r_moon = d->r_moon_km;
    a.inv_lambda_km = 1e3 / g->lambda_m;
    a.raan = g->d_orbit;
    a.m0 = g->d_orbit + g->n_sat;
    a.t_idx = g->d_tidx;
    a.ceil_uvw = (d->uvw_flags & VRS_UVW_CEIL) ? 1 : 0;
    // the host-side copies above are pageable; finish them before the vectors go away
    VRS_HIP(ctx, hipStreamSynchronize(st));
    return VRS_OK;
}

vrs_status vrs_geom_create(vrs_ctx *ctx, const vrs_geom_desc *d, vrs_geom **out) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!d || !out) return fail(ctx, VRS_E_ARG, "desc and out must be non-NULL");
    *out = nullptr;
    const vrs_orbit_desc &o = d->orbit;
    if (o.n_sat < 2 || o.n_sat > 4096) return fail(ctx, VRS_E_ARG, "n_sat must be in [2, 4096]");
    if (o.n_t < 1) return fail(ctx, VRS_E_ARG, "n_t must be >= 1");
    if (!(o.dt_s > 0) || !std::isfinite(o.dt_s)) return fail(ctx, VRS_E_ARG, "dt_s must be > 0");
    if (!(o.mu_km3_s2 > 0)) return fail(ctx, VRS_E_ARG, "mu must be > 0");
    if (!o.raan_rad || !o.m0_rad) return fail(ctx, VRS_E_ARG, "raan_rad and m0_rad are required");
    for (int k = 0; k < o.n_sat; ++k)
        if (!std::isfinite(o.raan_rad[k]) || !std::isfinite(o.m0_rad[k]))
            return fail(ctx, VRS_E_ARG, "orbit angles of satellite %d are not finite", k);
    if (!(d->r_moon_km > 0)) return fail(ctx, VRS_E_ARG, "r_moon_km must be > 0");
    if (!(o.a_km > d->r_moon_km * (1.0 + 1e-3)))
        return fail(ctx, VRS_E_ARG, "orbit radius %.3f km must exceed the Moon radius %.3f km",
                    o.a_km, d->r_moon_km);
    if (!(d->freq_hz > 0) || !std::isfinite(d->freq_hz))
        return fail(ctx, VRS_E_ARG, "freq_hz must be > 0");
    if (d->nside < 1 || d->nside > 4096 || (d->nside & (d->nside - 1)))
        return fail(ctx, VRS_E_ARG, "nside must be a power of two in [1, 4096]");
    if (d->n_base < 0) return fail(ctx, VRS_E_ARG, "n_base must be >= 0");
    if (d->n_base > 0) {
        if (!d->pairs) return fail(ctx, VRS_E_ARG, "pairs required when n_base > 0");
        for (int b = 0; b < d->n_base; ++b) {
            int pa = d->pairs[2 * b], pb = d->pairs[2 * b + 1];
            if (pa < 0 || pb < 0 || pa >= o.n_sat || pb >= o.n_sat || pa == pb)
                return fail(ctx, VRS_E_ARG, "baseline %d has invalid satellites (%d, %d)", b, pa, pb);
        }
    }
    const int64_t nb = d->n_base ? d->n_base : (int64_t)o.n_sat * (o.n_sat - 1) / 2;
    if (nb * (int64_t)o.n_t > (int64_t)1 << 31)
        return fail(ctx, VRS_E_ARG, "n_base * n_t = %lld exceeds 2^31", (long long)(nb * o.n_t));
    if ((int64_t)o.n_sat * o.n_t > (int64_t)1 << 31)
        return fail(ctx, VRS_E_ARG, "n_sat * n_t exceeds 2^31");
    if (d->uvw_flags & ~(VRS_UVW_CEIL | VRS_UVW_LAMBDA_INT))
        return fail(ctx, VRS_E_ARG, "unknown uvw_flags 0x%x", d->uvw_flags);
    if ((d->uvw_flags & VRS_UVW_LAMBDA_INT) && !(3e8 / d->freq_hz >= 1.0))
        return fail(ctx, VRS_E_ARG, "(int)(3e8 / freq_hz) is zero at %g Hz", d->freq_hz);
    if (o.sample_index)
        for (int j = 0; j < o.n_t; ++j)
            if (o.sample_index[j] > ((int64_t)1 << 52))
                return fail(ctx, VRS_E_ARG, "sample_index[%d] = %lld is too large", j,
                            (long long)o.sample_index[j]);
    DeviceGuard dg(ctx->device);
    vrs_geom *g = new vrs_geom;
    g->ctx = ctx;
    g->device = ctx->device;
    vrs_status s = geom_build(ctx, d, g);
    if (s == VRS_OK) s = vrs_orbit_uvw(ctx, g, nullptr, nullptr);
    if (s == VRS_OK) {
        VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *out = g;
    } else {
        vrs_geom_destroy(g);
    }
    return s;
}

vrs_status vrs_geom_create_uvw(vrs_ctx *ctx, const vrs_uvw_geom_desc *d, vrs_geom **out) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!d || !out) return fail(ctx, VRS_E_ARG, "desc and out must be non-NULL");
    *out = nullptr;
    if (d->n_rows < 1 || d->n_rows > ((int64_t)1 << 31))
        return fail(ctx, VRS_E_ARG, "n_rows must be in [1, 2^31]");
    if (!d->uvw) return fail(ctx, VRS_E_ARG, "uvw is required");
    for (int64_t i = 0; i < 3 * d->n_rows; ++i)
        if (!std::isfinite(d->uvw[i])) return fail(ctx, VRS_E_ARG, "uvw row %lld is not finite", (long long)(i / 3));
    if (d->lmn) {
        if (d->n_pix < 1 || d->n_pix > ((int64_t)1 << 30))
            return fail(ctx, VRS_E_ARG, "n_pix must be in [1, 2^30]");
        for (int64_t p = 0; p < d->n_pix; ++p) {
            const double *v = d->lmn + 3 * p;
            const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (!(std::fabs(r2 - 1.0) <= 1e-6))
                return fail(ctx, VRS_E_ARG, "lmn row %lld is not a unit vector", (long long)p);
        }
    } else if (d->nside < 1 || d->nside > 4096 || (d->nside & (d->nside - 1))) {
        return fail(ctx, VRS_E_ARG, "nside must be a power of two in [1, 4096] (or give lmn)");
    }
    DeviceGuard dg(ctx->device);
    vrs_geom *g = new vrs_geom;
    g->ctx = ctx;
    g->device = ctx->device;
    g->from_uvw = true;
    // one stand-in satellite step whose cap shades nothing; every row refers to it
    g->n_sat = 1;
    g->n_t = (int32_t)d->n_rows;
    g->n_base = 1;
    g->pairs = {0, 0};
    g->n_vis = d->n_rows;
    g->r_moon = 0.0;
    auto bail = [&](vrs_status s) { vrs_geom_destroy(g); return s; };
    if (vrs_status s = sky_build(ctx, g, d->nside, d->n_pix, d->lmn)) return bail(s);
    std::vector<double4> uvw(g->n_vis);
    for (int64_t v = 0; v < g->n_vis; ++v)
        uvw[v] = make_double4(d->uvw[3 * v], d->uvw[3 * v + 1], d->uvw[3 * v + 2], 0.0);
    const CapRec none = {0.f, 0.f, 0.f, 2.0f, 0.f, kGuardMin, 2.0f - kGuardMin, 2.0f + kGuardMin};
    const double4 origin = make_double4(0.0, 0.0, 0.0, 0.0);
    hipStream_t st = ctx->stream;
    auto hip = [&](hipError_t e, const char *what) -> vrs_status {
        if (e == hipSuccess) return VRS_OK;
        return fail(ctx, e == hipErrorOutOfMemory ? VRS_E_OOM : VRS_E_HIP, "%s: %s", what,
                    hipGetErrorString(e));
    };
    vrs_status s = VRS_OK;
    if (!s) s = hip(hipMalloc(&g->d_pos, sizeof(double4)), "hipMalloc");
    if (!s) s = hip(hipMalloc(&g->d_cap, sizeof(CapRec)), "hipMalloc");
    if (!s) s = hip(hipMalloc(&g->d_uvw64, sizeof(double4) * g->n_vis), "hipMalloc");
    if (!s) s = hip(hipMalloc(&g->d_viscap, sizeof(int2) * g->n_vis), "hipMalloc");
    if (!s) s = hip(hipMalloc(&g->d_pairs, sizeof(int2)), "hipMalloc");
    if (!s) s = hip(hipMemcpyAsync(g->d_pos, &origin, sizeof origin, hipMemcpyHostToDevice, st), "copy");
    if (!s) s = hip(hipMemcpyAsync(g->d_cap, &none, sizeof none, hipMemcpyHostToDevice, st), "copy");
    if (!s) s = hip(hipMemcpyAsync(g->d_uvw64, uvw.data(), sizeof(double4) * g->n_vis,
                                   hipMemcpyHostToDevice, st), "copy");
    if (!s) s = hip(hipMemsetAsync(g->d_viscap, 0, sizeof(int2) * g->n_vis, st), "memset");
    if (!s) s = hip(hipMemsetAsync(g->d_pairs, 0, sizeof(int2), st), "memset");
    if (!s) s = hip(hipStreamSynchronize(st), "sync");
    if (s) return bail(s);
    *out = g;
    return VRS_OK;
}

vrs_status vrs_geom_info_get(const vrs_geom *g, vrs_geom_info *out) {
    if (!g || !out) return fail(nullptr, VRS_E_ARG, "geom and out must be non-NULL");
    out->n_vis = g->n_vis;
    out->n_pix = g->n_pix;
    out->n_sat = g->n_sat;
    out->n_t = g->n_t;
    out->n_base = g->n_base;
    out->nside = g->nside;
    out->n_chunk = g->n_chunk;
    out->chunk_pixels = kChunk;
    out->lambda_m = g->lambda_m;
    return VRS_OK;
}

vrs_status vrs_orbit_uvw(vrs_ctx *ctx, vrs_geom *g, double *sat_pos, double *uvw) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!g || g->ctx != ctx) return fail(ctx, VRS_E_ARG, "geometry does not belong to this context");
    if (g->from_uvw && sat_pos)
        return fail(ctx, VRS_E_STATE, "a geometry built from uvw rows has no satellite positions");
    DeviceGuard dg(ctx->device);
    if (!g->from_uvw) {
        Timer tm(ctx, "orbit");
        VRS_HIP(ctx, launch_orbit(g->orb, g->dev(), g->d_pos, g->d_cap, g->d_uvw64, g->d_viscap,
                                  ctx->stream));
    }
    // strided exports (x, y, z) out of the padded 4-wide device records
    if (sat_pos)
        VRS_HIP(ctx, hipMemcpy2DAsync(sat_pos, 3 * sizeof(double), g->d_pos, sizeof(double4),
                                      3 * sizeof(double), (size_t)g->n_sat * g->n_t,
                                      hipMemcpyDefault, ctx->stream));
    if (uvw)
        VRS_HIP(ctx, hipMemcpy2DAsync(uvw, 3 * sizeof(double), g->d_uvw64, sizeof(double4),
                                      3 * sizeof(double), (size_t)g->n_vis, hipMemcpyDefault,
                                      ctx->stream));
    if (sat_pos || uvw) VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VRS_OK;
}

vrs_status vrs_geom_lmn(vrs_ctx *ctx, const vrs_geom *g, double *lmn) {
    if (!g || !lmn) return fail(ctx, VRS_E_ARG, "geom and lmn must be non-NULL");
    if (is_device_ptr(lmn)) {
        DeviceGuard dg(g->ctx->device);
        // on the context stream, finished before the host vector can change
        VRS_HIP(ctx, hipMemcpyAsync(lmn, g->lmn_ring.data(), sizeof(double) * g->lmn_ring.size(),
                                    hipMemcpyHostToDevice, g->ctx->stream));
        VRS_HIP(ctx, hipStreamSynchronize(g->ctx->stream));
    } else {
        std::memcpy(lmn, g->lmn_ring.data(), sizeof(double) * g->lmn_ring.size());
    }
    return VRS_OK;
}

// ---------------------------------------------------------------- shade
vrs_status vrs_shade(vrs_ctx *ctx, const vrs_geom *g, int32_t *count, uint32_t *hash,
                     uint32_t *bits, vrs_shade_stats *stats) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!g || g->ctx != ctx) return fail(ctx, VRS_E_ARG, "geometry does not belong to this context");
    const int64_t n_words = (g->n_pix + 31) / 32;
    if (bits && g->n_vis * n_words * 4 > ((int64_t)1 << 32))
        return fail(ctx, VRS_E_ARG, "shade bitmap of %lld bytes exceeds 4 GiB",
                    (long long)(g->n_vis * n_words * 4));
    DeviceGuard dg(ctx->device);
    Staged sc(ctx, count, sizeof(int32_t) * g->n_vis, true);
    Staged sh(ctx, hash, sizeof(uint32_t) * g->n_vis, true);
    Staged sb(ctx, bits, sizeof(uint32_t) * g->n_vis * n_words, true);
    if (vrs_status s = sc.prepare()) return s;
    if (vrs_status s = sh.prepare()) return s;
    if (vrs_status s = sb.prepare()) return s;
    if (bits) VRS_HIP(ctx, hipMemsetAsync(sb.dev, 0, sb.bytes, ctx->stream));
    VRS_HIP(ctx, hipMemsetAsync(ctx->d_tot, 0, 4 * sizeof(unsigned long long), ctx->stream));
    const GeomDev gd = g->dev();
    const int grid = persistent_grid(ctx, shade_kernel_ptr(), (g->n_vis + 3) / 4);
    {
        Timer tm(ctx, "shade");
        VRS_HIP(ctx, launch_shade(gd, (int32_t *)sc.dev, (uint32_t *)sh.dev, (uint32_t *)sb.dev,
                                  n_words, ctx->d_tot, grid, ctx->stream));
    }
    if (vrs_status s = sc.finish()) return s;
    if (vrs_status s = sh.finish()) return s;
    if (vrs_status s = sb.finish()) return s;
    if (stats) {
        unsigned long long tot[4];
        VRS_HIP(ctx, hipMemcpyAsync(tot, ctx->d_tot, sizeof tot, hipMemcpyDeviceToHost, ctx->stream));
        VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        stats->shaded_pairs = (int64_t)tot[0];
        stats->near_limb = (int64_t)tot[1];
        stats->fp64_tests = (int64_t)tot[2];
    }
    return VRS_OK;
}

// ---------------------------------------------------------------- vis-gen / recon
static vrs_status check_geom(vrs_ctx *ctx, const vrs_geom *g) {
    if (!g) return fail(ctx, VRS_E_ARG, "geometry is NULL");
    if (g->ctx != ctx) return fail(ctx, VRS_E_ARG, "geometry does not belong to this context");
    return VRS_OK;
}

// Steps per block of the per-satellite-factor matrix-core vis-gen, 0 where it
// does not apply: u v w must be the (unrounded) satellite position differences.
static int sat_factor_tb(const vrs_geom *g, const vrs_knobs *k) {
    if (!k->sat_factors || g->from_uvw || (g->uvw_flags & VRS_UVW_CEIL)) return 0;
    return mfma_sat_tb(g->n_sat, g->n_base);
}

// Satellite count of the per-satellite-factor matrix-core recon, 0 where it does
// not apply: factorisable u v w (as for vis-gen), 8 or 16 satellites, and the
// lexicographic pair list (its operand layout is built on it).
static int sat_recon_ns(const vrs_geom *g, const vrs_knobs *k) {
    if (!k->sat_factors || g->from_uvw || (g->uvw_flags & VRS_UVW_CEIL)) return 0;
    const int ns = mfma_recon_sat_ns(g->n_sat);
    if (ns == 0 || g->n_base != ns * (ns - 1) / 2) return 0;
    int b = 0;
    for (int a = 0; a < ns; ++a)
        for (int c = a + 1; c < ns; ++c, ++b)
            if (g->pairs[2 * b] != a || g->pairs[2 * b + 1] != c) return 0;
    return ns;
}

vrs_status vrs_visgen(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, int32_t n_maps,
                      const float *I, float *V) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (n_maps < 1 || n_maps > 65536) return fail(ctx, VRS_E_ARG, "n_maps must be in [1, 65536]");
    if (!I || !V) return fail(ctx, VRS_E_ARG, "I and V must be non-NULL");
    DeviceGuard dg(ctx->device);
    Staged sI(ctx, I, sizeof(float) * n_maps * g->n_pix, false);
    Staged sV(ctx, V, sizeof(float) * 2 * n_maps * g->n_vis, true);
    if (vrs_status s = sI.prepare()) return s;
    if (vrs_status s = sV.prepare()) return s;
    if (k->precision == VRS_MFMA_SPLIT16 && n_maps == 1 && sat_factor_tb(g, k) > 0 &&
        g->n_pad % 32 == 0) {
        // one map: each step's visibilities as entries of Gram matrices of the satellite factors
        const int64_t n_seg = gram_visgen_segments(g->n_pad, g->n_t);
        void *part, *small;
        if (vrs_status s = ws_get(ctx, 0, gram_visgen_part_bytes(g->n_t, n_seg), &part)) return s;
        if (vrs_status s = ws_get(ctx, 3, 256 + quad_recon_rows_bytes(g->n_t), &small)) return s;
        Timer tm(ctx, "mfma_visgen");
        VRS_HIP(ctx, launch_gram_visgen(g->dev(), phase_knobs(k), g->orb.inv_lambda_km, g->n_pix,
                                        (const float *)sI.dev, (unsigned int *)small,
                                        (char *)small + 256, (int32_t)n_seg, (float *)part,
                                        (float2 *)sV.dev, ctx->stream));
        tm.stop();
        return sV.finish();
    }
    if (k->precision == VRS_MFMA_SPLIT16) {
        void *a16, *small;
        if (vrs_status s = ws_get(ctx, 0, mfma_a16_bytes_visgen(n_maps, g->n_pad), &a16)) return s;
        if (vrs_status s = ws_get(ctx, 3, 8 * (size_t)n_maps + 256, &small)) return s;
        const int64_t groups = (n_maps + 63) / 64;
        const int tb = sat_factor_tb(g, k);
        if (tb > 0) {
            const int64_t n_tiles = groups * ((g->n_t + tb - 1) / tb);
            const int grid = persistent_grid(ctx, mfma_visgen_sat_kernel_ptr(tb), n_tiles,
                                             mfma_block_threads());
            Timer tm(ctx, "mfma_visgen");
            const bool star = sat_recon_ns(g, k) > 0 && mfma_visgen_sat_star(g->n_sat, g->n_base);
            VRS_HIP(ctx, launch_mfma_visgen_sat(g->dev(), phase_knobs(k), g->orb.inv_lambda_km,
                                                n_maps, g->n_pix, (const float *)sI.dev,
                                                (unsigned int *)small, (float *)small + n_maps,
                                                a16, (float2 *)sV.dev, star, grid, ctx->stream));
            tm.stop();
            return sV.finish();
        }
        const int64_t n_tiles = groups * ((g->n_vis + mfma_vis_per_block() - 1) / mfma_vis_per_block());
        const int grid = persistent_grid(ctx, mfma_visgen_kernel_ptr(), n_tiles, mfma_block_threads());
        Timer tm(ctx, "mfma_visgen");
        VRS_HIP(ctx, launch_mfma_visgen(g->dev(), phase_knobs(k), n_maps, g->n_pix,
                                        (const float *)sI.dev, (unsigned int *)small,
                                        (float *)small + n_maps, a16, (float2 *)sV.dev, grid,
                                        ctx->stream));
        tm.stop();
        return sV.finish();
    }
    void *ich;
    if (vrs_status s = ws_get(ctx, 0, sizeof(float) * n_maps * g->n_pad, &ich)) return s;
    const GeomDev gd = g->dev();
    VRS_HIP(ctx, launch_to_chunk_order(gd, n_maps, g->n_pix, (const float *)sI.dev, (float *)ich,
                                       ctx->stream));
    const bool f64 = k->precision == VRS_MIXED_F64PHASE;
    const bool pair = k->antipodal && g->n_half > 0;
    int vt = visgen_tile_vis();
    {
        auto it = ctx->tuned_vt.find(std::make_tuple(g->n_vis, g->n_pix, n_maps, tune_mode(k)));
        if (it != ctx->tuned_vt.end()) vt = it->second;
    }
    const int64_t n_tiles = ((g->n_vis + vt - 1) / vt) * n_maps;
    const int grid = persistent_grid(ctx, visgen_kernel_ptr(f64, vt, pair), n_tiles);
    {
        Timer tm(ctx, "visgen");
        VRS_HIP(ctx, launch_visgen(gd, phase_knobs(k), f64, vt, pair, n_maps, (const float *)ich,
                                   (float2 *)sV.dev, grid, ctx->stream));
    }
    if (vrs_status s = sV.finish()) return s;
    return VRS_OK;
}

vrs_status vrs_autotune_visgen(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k,
                               int32_t n_maps, const float *I, int32_t *tile_out, double *ms_out) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (k->precision == VRS_MFMA_SPLIT16)
        return fail(ctx, VRS_E_ARG, "the matrix-core path has fixed tiles");
    if (n_maps < 1 || n_maps > 65536 || !I) return fail(ctx, VRS_E_ARG, "bad n_maps or I");
    DeviceGuard dg(ctx->device);
    void *V;
    if (vrs_status s = ws_get(ctx, 2, sizeof(float) * 2 * n_maps * g->n_vis, &V)) return s;
    const auto key = std::make_tuple(g->n_vis, g->n_pix, n_maps, tune_mode(k));
    hipEvent_t e0, e1;
    VRS_HIP(ctx, hipEventCreate(&e0));
    VRS_HIP(ctx, hipEventCreate(&e1));
    double best = 1e300;
    int best_vt = visgen_tile_vis();
    vrs_status st = VRS_OK;
    // the antipodal-pair kernel holds four sums per visibility: 32-visibility
    // tiles spill, so it is tuned over 8 and 16
    const bool pair = k->antipodal && g->n_half > 0;
    const std::vector<int> tiles = pair ? std::vector<int>{8, 16} : std::vector<int>{8, 16, 32};
    for (int rep = 0; rep < 2 && st == VRS_OK; ++rep)     // first round warms up
        for (int vt : tiles) {
            ctx->tuned_vt[key] = vt;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) { st = VRS_E_HIP; break; }
            st = vrs_visgen(ctx, g, k, n_maps, I, (float *)V);
            if (st != VRS_OK) break;
            float ms = 0;
            if (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { st = VRS_E_HIP; break; }
            if (rep == 1 && ms < best) { best = ms; best_vt = vt; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != VRS_OK) {
        ctx->tuned_vt.erase(key);
        return st == VRS_E_HIP ? fail(ctx, st, "event timing failed during autotune") : st;
    }
    ctx->tuned_vt[key] = best_vt;
    if (tile_out) *tile_out = best_vt;
    if (ms_out) *ms_out = best;
    return VRS_OK;
}

// Segments of the visibility range of the VALU recon: enough work items to fill
// the chip, a multiple of 8 so that the blocks on each XCD stream their own
// segments; independent of n_maps, so that each map's sums are grouped the same way.
static int64_t recon_segments(const vrs_geom *g) {
    int64_t n_seg = (4096 + g->n_chunk - 1) / g->n_chunk;
    n_seg = (n_seg + 7) / 8 * 8;
    while (n_seg > 8 && g->n_vis / n_seg < 1024) n_seg -= 8;
    return std::min<int64_t>(n_seg, 1024);
}

static vrs_status recon_common(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k,
                               int32_t n_maps, const float *V, const float *w, float *F,
                               double kappa, const char *timer_name) {
    DeviceGuard dg(ctx->device);
    Staged sV(ctx, V, sizeof(float) * 2 * n_maps * g->n_vis, false);
    Staged sw(ctx, w, sizeof(float) * g->n_vis, false);
    Staged sF(ctx, F, sizeof(float) * n_maps * g->n_pix, true);
    if (vrs_status s = sV.prepare()) return s;
    if (vrs_status s = sw.prepare()) return s;
    if (vrs_status s = sF.prepare()) return s;
    if (k->precision == VRS_MFMA_SPLIT16 && n_maps == 1 && sat_factor_tb(g, k) > 0 &&
        g->n_pad % 64 == 0) {
        // one map: the pairs of a step as a Hermitian form in the satellite factors
        const int64_t n_seg = quad_recon_segments(g->n_pad, g->n_t);
        const size_t small_b = 256;
        void *wop, *small, *part;
        if (vrs_status s = ws_get(ctx, 2, quad_recon_wop_bytes(g->n_t), &wop)) return s;
        if (vrs_status s = ws_get(ctx, 3, small_b + quad_recon_rows_bytes(g->n_t), &small)) return s;
        if (vrs_status s = ws_get(ctx, 1, sizeof(float) * n_seg * g->n_pad, &part)) return s;
        const GeomDev gd = g->dev();
        {
            Timer tm(ctx, "mfma_recon");
            VRS_HIP(ctx, launch_quad_recon(gd, phase_knobs(k), g->orb.inv_lambda_km,
                                           (const float2 *)sV.dev, (const float *)sw.dev,
                                           (float)kappa, (unsigned int *)small, (float *)small + 1,
                                           wop, (char *)small + small_b, (int32_t)n_seg,
                                           (float *)part, ctx->stream));
        }
        VRS_HIP(ctx, launch_recon_finalize(gd, 1, (int32_t)n_seg, g->n_pix, (const float *)part,
                                           (float *)sF.dev, ctx->stream));
        return sF.finish();
    }
    if (k->precision == VRS_MFMA_SPLIT16 && sat_recon_ns(g, k) > 0) {
        const int64_t n_slots = mfma_recon_sat_slots(g->n_sat, g->n_t);
        const int64_t groups = (n_maps + 63) / 64;
        const int64_t ptiles = (g->n_pad + mfma_recon_sat_pix_per_block() - 1) / mfma_recon_sat_pix_per_block();
        const int64_t n_tp = ((int64_t)g->n_t + 1) / 2;
        int64_t n_seg = (2048 + ptiles * groups - 1) / (ptiles * groups);
        n_seg = std::max<int64_t>(1, std::min<int64_t>(n_seg, n_tp));
        if (n_seg > 8) n_seg = std::min<int64_t>((n_seg + 7) / 8 * 8, n_tp);
        const size_t small_b = (8 * (size_t)n_maps + 256 + 255) / 256 * 256;
        const size_t spos_b = sizeof(double) * 4 * (size_t)g->n_sat * g->n_t;
        void *a16, *small, *part;
        if (vrs_status s = ws_get(ctx, 2, (size_t)groups * 4 * 64 * n_slots * 2, &a16)) return s;
        if (vrs_status s = ws_get(ctx, 3, small_b + spos_b, &small)) return s;
        if (vrs_status s = ws_get(ctx, 1, sizeof(float) * n_seg * n_maps * g->n_pad, &part)) return s;
        const int ns = sat_recon_ns(g, k);
        const int grid = persistent_grid(ctx, mfma_recon_sat_kernel_ptr(ns), ptiles * groups * n_seg,
                                         mfma_recon_sat_threads());
        const GeomDev gd = g->dev();
        {
            Timer tm(ctx, "mfma_recon");
            VRS_HIP(ctx, launch_mfma_recon_sat(gd, phase_knobs(k), g->orb.inv_lambda_km, n_maps,
                                               (const float2 *)sV.dev, (const float *)sw.dev,
                                               (float)kappa, n_slots, (unsigned int *)small,
                                               (float *)small + n_maps, a16, (char *)small + small_b,
                                               (int32_t)n_seg, (float *)part, grid, ctx->stream));
        }
        VRS_HIP(ctx, launch_recon_finalize(gd, n_maps, (int32_t)n_seg, g->n_pix,
                                           (const float *)part, (float *)sF.dev, ctx->stream));
        return sF.finish();
    }
    if (k->precision == VRS_MFMA_SPLIT16) {
        int64_t n_vis_pad;
        mfma_vis_pad(g->n_vis, &n_vis_pad);
        const int64_t groups = (n_maps + 63) / 64;
        const int64_t ptiles = (g->n_pad + mfma_pix_per_block() - 1) / mfma_pix_per_block();
        int64_t n_seg = (2048 + ptiles * groups - 1) / (ptiles * groups);
        n_seg = std::max<int64_t>(1, std::min<int64_t>(n_seg, n_vis_pad / 1024));
        if (n_seg > 8) n_seg = (n_seg + 7) / 8 * 8;
        void *a16, *small, *part;
        if (vrs_status s = ws_get(ctx, 2, mfma_a16_bytes_recon(n_maps, n_vis_pad), &a16)) return s;
        if (vrs_status s = ws_get(ctx, 3, 8 * (size_t)n_maps + 256, &small)) return s;
        if (vrs_status s = ws_get(ctx, 1, sizeof(float) * n_seg * n_maps * g->n_pad, &part)) return s;
        const int grid = persistent_grid(ctx, mfma_recon_kernel_ptr(), ptiles * groups * n_seg,
                                         mfma_recon_block_threads());
        const GeomDev gd = g->dev();
        {
            Timer tm(ctx, "mfma_recon");
            VRS_HIP(ctx, launch_mfma_recon(gd, phase_knobs(k), n_maps, (const float2 *)sV.dev,
                                           (const float *)sw.dev, (float)kappa, n_vis_pad,
                                           (unsigned int *)small, (float *)small + n_maps, a16,
                                           (int32_t)n_seg, (float *)part, grid, ctx->stream));
        }
        VRS_HIP(ctx, launch_recon_finalize(gd, n_maps, (int32_t)n_seg, g->n_pix,
                                           (const float *)part, (float *)sF.dev, ctx->stream));
        return sF.finish();
    }
    const int64_t units = (int64_t)n_maps * g->n_chunk;
    const int64_t n_seg = recon_segments(g);
    void *part;
    if (vrs_status s = ws_get(ctx, 1, sizeof(float) * n_seg * n_maps * g->n_pad, &part)) return s;
    const GeomDev gd = g->dev();
    const bool f64 = k->precision == VRS_MIXED_F64PHASE;
    const int grid = persistent_grid(ctx, recon_kernel_ptr(f64), units * n_seg);
    {
        Timer tm(ctx, timer_name);
        VRS_HIP(ctx, launch_recon(gd, phase_knobs(k), f64, n_maps, (const float2 *)sV.dev,
                                  (const float *)sw.dev, (float)kappa, (int32_t)n_seg,
                                  (float *)part, ctx->d_tot + 4, grid, ctx->stream));
    }
    VRS_HIP(ctx, launch_recon_finalize(gd, n_maps, (int32_t)n_seg, g->n_pix, (const float *)part,
                                       (float *)sF.dev, ctx->stream));
    if (vrs_status s = sF.finish()) return s;
    return VRS_OK;
}

vrs_status vrs_recon(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, int32_t n_maps,
                     const float *V, const float *w, float *F) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (n_maps < 1 || n_maps > 65536) return fail(ctx, VRS_E_ARG, "n_maps must be in [1, 65536]");
    if (!V || !F) return fail(ctx, VRS_E_ARG, "V and F must be non-NULL");
    const double kappa = k->kappa > 0 ? k->kappa : 1.0 / (double)g->n_vis;
    return recon_common(ctx, g, k, n_maps, V, w, F, kappa, "recon");
}

vrs_status vrs_adjoint(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, const float *resid,
                       float *grad) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (!resid || !grad) return fail(ctx, VRS_E_ARG, "resid and grad must be non-NULL");
    vrs_knobs kk = *k;
    if (kk.precision == VRS_MFMA_SPLIT16) kk.precision = VRS_MIXED_F64PHASE;
    return recon_common(ctx, g, &kk, 1, resid, nullptr, grad, 2.0, "adjoint");
}

vrs_status vrs_recon_noalloc(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, int32_t n_maps,
                             const float *V, const float *w, float *F) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (k->precision == VRS_MFMA_SPLIT16)
        return fail(ctx, VRS_E_ARG, "vrs_recon_noalloc runs VRS_PLAIN_F32 or VRS_MIXED_F64PHASE");
    if (n_maps < 1 || n_maps > 65536) return fail(ctx, VRS_E_ARG, "n_maps must be in [1, 65536]");
    if (!V || !F) return fail(ctx, VRS_E_ARG, "V and F must be non-NULL");
    if (!ctx->noalloc_scratch) return fail(ctx, VRS_E_STATE, "no recon scratch on this device");
    DeviceGuard dg(ctx->device);
    // host arrays would need staging buffers
    if (!is_device_ptr(V) || !is_device_ptr(F) || (w && !is_device_ptr(w)))
        return fail(ctx, VRS_E_ARG, "vrs_recon_noalloc takes device arrays only (V, w, F)");
    const int64_t n_seg = recon_segments(g);
    const uint64_t need = (uint64_t)n_seg * (uint64_t)n_maps * (uint64_t)g->n_pad * sizeof(float);
    if (need > (uint64_t)(VRS_NOALLOC_SCRATCH_BYTES - kNoallocQueueBytes))
        return fail(ctx, VRS_E_ARG,
                    "recon of %d map(s) over %lld padded pixels needs %llu bytes of partial "
                    "images; vrs_recon_noalloc holds %llu (use vrs_recon or fewer maps)",
                    n_maps, (long long)g->n_pad, (unsigned long long)need,
                    (unsigned long long)(VRS_NOALLOC_SCRATCH_BYTES - kNoallocQueueBytes));
    const double kappa = k->kappa > 0 ? k->kappa : 1.0 / (double)g->n_vis;
    const GeomDev gd = g->dev();
    const bool f64 = k->precision == VRS_MIXED_F64PHASE;
    const int grid = persistent_grid(ctx, recon_kernel_ptr(f64), (int64_t)n_maps * g->n_chunk * n_seg);
    // every (slot, chunk) item writes its whole partial-image block: nothing to clear first
    unsigned long long *queue = (unsigned long long *)((char *)ctx->noalloc_scratch +
                                                       VRS_NOALLOC_SCRATCH_BYTES - kNoallocQueueBytes);
    VRS_HIP(ctx, launch_recon(gd, phase_knobs(k), f64, n_maps, (const float2 *)V, w, (float)kappa,
                              (int32_t)n_seg, ctx->noalloc_scratch, queue, grid, ctx->stream));
    VRS_HIP(ctx, launch_recon_finalize(gd, n_maps, (int32_t)n_seg, g->n_pix,
                                       ctx->noalloc_scratch, F, ctx->stream));
    return VRS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- ragged batches
struct vrs_batch {
    vrs_ctx *ctx = nullptr;
    const vrs_geom *g = nullptr;
    int device = 0;
    int32_t n_ch = 0, precision = 0;
    int64_t n_vis = 0, n_pad_vis = 0, n_slots = 0;
    int vt = 16, seg = 16384;
    double tune_v = 0, tune_r = 0;
    std::vector<int64_t> offsets, kept;   // [n_ch + 1], [n_ch]
    std::vector<int32_t> sample_bt;       // [n_pad_vis]
    std::vector<double> inv_lambda;
    // device
    int2 *d_vis_sat = nullptr;
    int32_t *d_tile_map = nullptr;
    double *d_inv_lambda = nullptr;
    int4 *d_slots = nullptr;
    int32_t *d_map_slot0 = nullptr;
    float *d_kappa = nullptr;
    unsigned long long *d_queue = nullptr;  // recon work counter (batch stream)
    float *d_I = nullptr, *d_Ich = nullptr, *d_F = nullptr, *d_part = nullptr;
    float2 *d_V = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    // knobs the graph was captured with
    double cap_s = 0, cap_kappa = -1;
    int32_t cap_wfac = -1, cap_pair = -1;
    float cap_tocc = 0;
    GeomDev dev() const {
        GeomDev d = g->dev();
        d.n_vis = n_pad_vis;
        d.vis_sat = d_vis_sat;
        d.tile_map = d_tile_map;
        d.inv_lambda = d_inv_lambda;
        d.slots = d_slots;
        d.map_slot0 = d_map_slot0;
        d.kappa_map = d_kappa;
        d.n_slots = n_slots;
        return d;
    }
};

namespace {

// Satellites a and b see each other: the segment between them misses the closed sphere.
bool line_of_sight(const double *ra, const double *rb, double R) {
    const double dx = rb[0] - ra[0], dy = rb[1] - ra[1], dz = rb[2] - ra[2];
    const double dd = dx * dx + dy * dy + dz * dz;
    double t = dd > 0 ? -(ra[0] * dx + ra[1] * dy + ra[2] * dz) / dd : 0.0;
    t = std::min(1.0, std::max(0.0, t));
    const double px = ra[0] + t * dx, py = ra[1] + t * dy, pz = ra[2] + t * dz;
    return px * px + py * py + pz * pz > R * R;
}

vrs_status batch_slots(vrs_batch *b, int seg, std::vector<int4> &slots, std::vector<int32_t> &slot0) {
    slots.clear();
    slot0.assign(b->n_ch + 1, 0);
    for (int c = 0; c < b->n_ch; ++c) {
        slot0[c] = (int32_t)slots.size();
        const int64_t v0 = b->offsets[c], v1 = b->offsets[c + 1];
        for (int64_t v = v0; v < v1; v += seg) {
            int4 sl;
            sl.x = c;
            sl.y = (int32_t)v;
            sl.z = (int32_t)std::min<int64_t>(v1, v + seg);
            sl.w = 0;
            slots.push_back(sl);
        }
    }
    slot0[b->n_ch] = (int32_t)slots.size();
    return VRS_OK;
}

// Upload slots for range length `seg` and size the partial-image buffer.
vrs_status batch_set_seg(vrs_batch *b, int seg) {
    vrs_ctx *ctx = b->ctx;
    std::vector<int4> slots;
    std::vector<int32_t> slot0;
    batch_slots(b, seg, slots, slot0);
    VRS_HIP(ctx, hipStreamSynchronize(b->stream));
    if (b->d_slots) VRS_HIP(ctx, hipFree(b->d_slots));
    if (b->d_part) VRS_HIP(ctx, hipFree(b->d_part));
    b->d_slots = nullptr;
    b->d_part = nullptr;
    VRS_HIP(ctx, hipMalloc(&b->d_slots, sizeof(int4) * std::max<size_t>(slots.size(), 1)));
    VRS_HIP(ctx, hipMalloc(&b->d_part, sizeof(float) * std::max<size_t>(slots.size(), 1) * b->g->n_pad));
    // on the batch stream (the one its graph runs on); the host vectors end here
    VRS_HIP(ctx, hipMemcpyAsync(b->d_slots, slots.data(), sizeof(int4) * slots.size(),
                                hipMemcpyHostToDevice, b->stream));
    VRS_HIP(ctx, hipMemcpyAsync(b->d_map_slot0, slot0.data(), sizeof(int32_t) * slot0.size(),
                                hipMemcpyHostToDevice, b->stream));
    VRS_HIP(ctx, hipStreamSynchronize(b->stream));
    b->n_slots = (int64_t)slots.size();
    b->seg = seg;
    return VRS_OK;
}

// The kernels of one batch step, launched on the batch stream (also the capture body).
vrs_status batch_launch(vrs_batch *b, const PhaseKnobs &pk, bool skip_recon = false,
                        bool skip_visgen = false) {
    vrs_ctx *ctx = b->ctx;
    const GeomDev gd = b->dev();
    const bool f64 = b->precision == VRS_MIXED_F64PHASE;
    if (!skip_visgen) {
        VRS_HIP(ctx, launch_to_chunk_order(gd, b->n_ch, b->g->n_pix, b->d_I, b->d_Ich, b->stream));
        const bool pair = pk.pair && gd.n_half > 0;
        const int grid = persistent_grid(ctx, visgen_kernel_ptr(f64, b->vt, pair), b->n_pad_vis / b->vt);
        VRS_HIP(ctx, launch_visgen(gd, pk, f64, b->vt, pair, b->n_ch, b->d_Ich, b->d_V, grid,
                                   b->stream));
    }
    if (!skip_recon) {
        const int grid = persistent_grid(ctx, recon_kernel_ptr(f64), b->n_slots * b->g->n_chunk);
        VRS_HIP(ctx, launch_recon(gd, pk, f64, b->n_ch, b->d_V, nullptr, 0.0f, 1, b->d_part, b->d_queue, grid,
                                  b->stream));
        VRS_HIP(ctx, launch_recon_finalize(gd, b->n_ch, 1, b->g->n_pix, b->d_part, b->d_F, b->stream));
    }
    return VRS_OK;
}

vrs_status batch_time(vrs_batch *b, const PhaseKnobs &pk, bool visgen, double *ms) {
    vrs_ctx *ctx = b->ctx;
    VRS_HIP(ctx, hipEventRecord(b->ev_t0, b->stream));
    if (vrs_status s = batch_launch(b, pk, visgen, !visgen)) return s;
    VRS_HIP(ctx, hipEventRecord(b->ev_t1, b->stream));
    VRS_HIP(ctx, hipEventSynchronize(b->ev_t1));
    float f;
    VRS_HIP(ctx, hipEventElapsedTime(&f, b->ev_t0, b->ev_t1));
    *ms = f;
    return VRS_OK;
}

vrs_status batch_capture(vrs_batch *b, const PhaseKnobs &pk) {
    vrs_ctx *ctx = b->ctx;
    if (b->exec) {
        VRS_HIP(ctx, hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        VRS_HIP(ctx, hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    VRS_HIP(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    vrs_status st = batch_launch(b, pk);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (st != VRS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return fail(ctx, VRS_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    b->graph = graph;
    VRS_HIP(ctx, hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
    b->cap_s = pk.s;
    b->cap_wfac = pk.wfac;
    b->cap_tocc = pk.t_occ;
    b->cap_pair = pk.pair;
    return VRS_OK;
}

}  // namespace

extern "C" {

void vrs_batch_destroy(vrs_batch *b) {
    if (!b) return;
    DeviceGuard dg(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void *ptrs[] = {b->d_vis_sat, b->d_tile_map, b->d_inv_lambda, b->d_slots, b->d_map_slot0,
                    b->d_kappa, b->d_I, b->d_Ich, b->d_F, b->d_part, b->d_V, b->d_queue};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    hipEvent_t evs[] = {b->ev_in, b->ev_out, b->ev_t0, b->ev_t1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

vrs_status vrs_batch_create(vrs_ctx *ctx, const vrs_geom *g, const vrs_batch_desc *d,
                            vrs_batch **out) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (!d || !out) return fail(ctx, VRS_E_ARG, "desc and out must be non-NULL");
    *out = nullptr;
    if (g->from_uvw)
        return fail(ctx, VRS_E_STATE, "a batch needs satellite positions (uvw-row geometry)");
    if (d->n_ch < 1 || d->n_ch > 4096) return fail(ctx, VRS_E_ARG, "n_ch must be in [1, 4096]");
    if (!d->freq_hz || !d->n_t) return fail(ctx, VRS_E_ARG, "freq_hz and n_t are required");
    if (d->precision != VRS_PLAIN_F32 && d->precision != VRS_MIXED_F64PHASE)
        return fail(ctx, VRS_E_ARG, "batches run VRS_PLAIN_F32 or VRS_MIXED_F64PHASE");
    for (int c = 0; c < d->n_ch; ++c) {
        if (!(d->freq_hz[c] > 0) || !std::isfinite(d->freq_hz[c]))
            return fail(ctx, VRS_E_ARG, "freq_hz[%d] must be > 0", c);
        if (d->n_t[c] < 1 || d->n_t[c] > g->n_t)
            return fail(ctx, VRS_E_ARG, "n_t[%d] = %d outside [1, %d]", c, d->n_t[c], g->n_t);
    }
    DeviceGuard dg(ctx->device);
    // positions for the line-of-sight test (FP64, exported from stage 1)
    std::vector<double> pos((size_t)g->n_sat * g->n_t * 3);
    VRS_HIP(ctx, hipMemcpy2DAsync(pos.data(), 3 * sizeof(double), g->d_pos, sizeof(double4),
                                  3 * sizeof(double), (size_t)g->n_sat * g->n_t,
                                  hipMemcpyDeviceToHost, ctx->stream));
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    vrs_batch *b = new vrs_batch;
    b->ctx = ctx;
    b->g = g;
    b->device = ctx->device;
    b->n_ch = d->n_ch;
    b->precision = d->precision;
    const int pad = list_pad();
    std::vector<int2> vis_sat;
    std::vector<int32_t> tile_map;
    b->offsets.push_back(0);
    for (int c = 0; c < d->n_ch; ++c) {
        int64_t kept = 0;
        for (int bl = 0; bl < g->n_base; ++bl) {
            const int pa = g->pairs[2 * bl], pb = g->pairs[2 * bl + 1];
            for (int t = 0; t < d->n_t[c]; ++t) {
                const int ia = pa * g->n_t + t, ib = pb * g->n_t + t;
                if (d->compact && !line_of_sight(&pos[3 * (size_t)ia], &pos[3 * (size_t)ib], g->r_moon))
                    continue;
                vis_sat.push_back(make_int2(ia, ib));
                b->sample_bt.push_back(bl * g->n_t + t);
                ++kept;
            }
        }
        while (vis_sat.size() % pad) {
            vis_sat.push_back(make_int2(-1, -1));
            b->sample_bt.push_back(-1);
        }
        b->kept.push_back(kept);
        b->offsets.push_back((int64_t)vis_sat.size());
        b->n_vis += kept;
        b->inv_lambda.push_back(1e3 / (kC / d->freq_hz[c]));
    }
    for (int c = 0; c < d->n_ch; ++c)
        for (int64_t v = b->offsets[c]; v < b->offsets[c + 1]; v += pad) tile_map.push_back(c);
    b->n_pad_vis = (int64_t)vis_sat.size();
    if (b->n_pad_vis == 0 || b->n_pad_vis > ((int64_t)1 << 31)) {
        vrs_batch_destroy(b);
        return fail(ctx, VRS_E_ARG, "batch has %lld samples (must be in [1, 2^31])",
                    (long long)b->n_pad_vis);
    }
    auto cleanup = [&](vrs_status s) { vrs_batch_destroy(b); return s; };
#define VRS_BHIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return cleanup(fail(ctx, e_ == hipErrorOutOfMemory ? VRS_E_OOM : VRS_E_HIP,       \
                                "%s failed: %s", #call, hipGetErrorString(e_)));              \
    } while (0)
    VRS_BHIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    VRS_BHIP(hipEventCreate(&b->ev_in));
    VRS_BHIP(hipEventCreate(&b->ev_out));
    VRS_BHIP(hipEventCreate(&b->ev_t0));
    VRS_BHIP(hipEventCreate(&b->ev_t1));
    VRS_BHIP(hipMalloc(&b->d_vis_sat, sizeof(int2) * b->n_pad_vis));
    VRS_BHIP(hipMalloc(&b->d_tile_map, sizeof(int32_t) * tile_map.size()));
    VRS_BHIP(hipMalloc(&b->d_inv_lambda, sizeof(double) * b->n_ch));
    VRS_BHIP(hipMalloc(&b->d_map_slot0, sizeof(int32_t) * (b->n_ch + 1)));
    VRS_BHIP(hipMalloc(&b->d_kappa, sizeof(float) * b->n_ch));
    VRS_BHIP(hipMalloc(&b->d_queue, sizeof(unsigned long long)));
    VRS_BHIP(hipMalloc(&b->d_I, sizeof(float) * b->n_ch * g->n_pix));
    VRS_BHIP(hipMalloc(&b->d_Ich, sizeof(float) * b->n_ch * g->n_pad));
    VRS_BHIP(hipMalloc(&b->d_F, sizeof(float) * b->n_ch * g->n_pix));
    VRS_BHIP(hipMalloc(&b->d_V, sizeof(float2) * b->n_pad_vis));
    // set-up on the batch stream (not captured: the graph is built later), finished
    // before the host vectors go away
    VRS_BHIP(hipMemsetAsync(b->d_V, 0, sizeof(float2) * b->n_pad_vis, b->stream));  // padding stays zero
    VRS_BHIP(hipMemsetAsync(b->d_I, 0, sizeof(float) * b->n_ch * g->n_pix, b->stream));
    VRS_BHIP(hipMemcpyAsync(b->d_vis_sat, vis_sat.data(), sizeof(int2) * vis_sat.size(),
                            hipMemcpyHostToDevice, b->stream));
    VRS_BHIP(hipMemcpyAsync(b->d_tile_map, tile_map.data(), sizeof(int32_t) * tile_map.size(),
                            hipMemcpyHostToDevice, b->stream));
    VRS_BHIP(hipMemcpyAsync(b->d_inv_lambda, b->inv_lambda.data(), sizeof(double) * b->n_ch,
                            hipMemcpyHostToDevice, b->stream));
    VRS_BHIP(hipStreamSynchronize(b->stream));
    if (vrs_status s = batch_set_seg(b, 16384)) return cleanup(s);

    // autotune: vis-gen tile (8, 16, 32 visibilities) and recon range length
    const auto key = std::make_tuple(b->n_pad_vis, g->n_pix, b->n_ch, b->precision);
    auto hit = ctx->tuned.find(key);
    if (hit != ctx->tuned.end()) {
        b->vt = hit->second.first;
        if (vrs_status s = batch_set_seg(b, hit->second.second)) return cleanup(s);
    } else if (d->autotune) {
        PhaseKnobs pk;
        pk.s = 1.0;
        pk.wfac = 0;
        pk.t_occ = 240.0f;
        pk.pair = 1;
        double best = 1e300, ms = 0;
        int best_vt = 16;
        if (vrs_status s = batch_time(b, pk, true, &ms)) return cleanup(s);  // warm-up
        const std::vector<int> tiles =
            g->n_half > 0 ? std::vector<int>{8, 16} : std::vector<int>{8, 16, 32};
        for (int vt : tiles) {
            b->vt = vt;
            if (vrs_status s = batch_time(b, pk, true, &ms)) return cleanup(s);
            if (ms < best) { best = ms; best_vt = vt; }
        }
        b->vt = best_vt;
        b->tune_v = best;
        best = 1e300;
        int best_seg = 16384;
        for (int seg : {4096, 16384, 65536}) {
            if (vrs_status s = batch_set_seg(b, seg)) return cleanup(s);
            if (vrs_status s = batch_time(b, pk, false, &ms)) return cleanup(s);
            if (ms < best) { best = ms; best_seg = seg; }
        }
        b->tune_r = best;
        if (vrs_status s = batch_set_seg(b, best_seg)) return cleanup(s);
        ctx->tuned[key] = std::make_pair(b->vt, b->seg);
    }
#undef VRS_BHIP
    *out = b;
    return VRS_OK;
}

vrs_status vrs_batch_info_get(const vrs_batch *b, vrs_batch_info *out) {
    if (!b || !out) return fail(nullptr, VRS_E_ARG, "batch and out must be non-NULL");
    out->n_vis_total = b->n_vis;
    out->n_vis_padded = b->n_pad_vis;
    out->n_slots = b->n_slots;
    out->vis_tile = b->vt;
    out->recon_seg = b->seg;
    out->tune_visgen_ms = b->tune_v;
    out->tune_recon_ms = b->tune_r;
    return VRS_OK;
}

vrs_status vrs_batch_layout(const vrs_batch *b, int64_t *offsets, int64_t *kept, int32_t *sample_bt) {
    if (!b) return fail(nullptr, VRS_E_ARG, "batch is NULL");
    if (offsets) std::memcpy(offsets, b->offsets.data(), sizeof(int64_t) * b->offsets.size());
    if (kept) std::memcpy(kept, b->kept.data(), sizeof(int64_t) * b->kept.size());
    if (sample_bt) std::memcpy(sample_bt, b->sample_bt.data(), sizeof(int32_t) * b->sample_bt.size());
    return VRS_OK;
}

vrs_status vrs_batch_run(vrs_ctx *ctx, vrs_batch *b, const vrs_knobs *k, const float *I, float *V,
                         float *F) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!b || b->ctx != ctx) return fail(ctx, VRS_E_ARG, "batch does not belong to this context");
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (!I) return fail(ctx, VRS_E_ARG, "I must be non-NULL");
    if (k->precision != b->precision)
        return fail(ctx, VRS_E_ARG, "knobs precision %d differs from the batch's %d", k->precision,
                    b->precision);
    DeviceGuard dg(ctx->device);
    const vrs_geom *g = b->g;
    // per-channel normalisation: kappa, or 1 / kept samples
    std::vector<float> kap(b->n_ch);
    for (int c = 0; c < b->n_ch; ++c)
        kap[c] = (float)(k->kappa > 0 ? k->kappa : (b->kept[c] > 0 ? 1.0 / (double)b->kept[c] : 0.0));
    const PhaseKnobs pk = phase_knobs(k);
    if (!b->exec || pk.s != b->cap_s || pk.wfac != b->cap_wfac || pk.t_occ != b->cap_tocc ||
        pk.pair != b->cap_pair) {
        if (vrs_status s = batch_capture(b, pk)) return s;
    }
    // order after the caller's stream, run on the batch stream, order the caller after it
    VRS_HIP(ctx, hipEventRecord(b->ev_in, ctx->stream));
    VRS_HIP(ctx, hipStreamWaitEvent(b->stream, b->ev_in, 0));
    VRS_HIP(ctx, hipMemcpyAsync(b->d_kappa, kap.data(), sizeof(float) * b->n_ch,
                                hipMemcpyHostToDevice, b->stream));
    VRS_HIP(ctx, hipMemcpyAsync(b->d_I, I, sizeof(float) * b->n_ch * g->n_pix, hipMemcpyDefault,
                                b->stream));
    if (ctx->timing) {
        auto it = ctx->events.find("batch");
        if (it == ctx->events.end()) {
            hipEvent_t e0, e1;
            VRS_HIP(ctx, hipEventCreate(&e0));
            VRS_HIP(ctx, hipEventCreate(&e1));
            it = ctx->events.emplace("batch", std::make_pair(e0, e1)).first;
        }
        VRS_HIP(ctx, hipEventRecord(it->second.first, b->stream));
        VRS_HIP(ctx, hipGraphLaunch(b->exec, b->stream));
        VRS_HIP(ctx, hipEventRecord(it->second.second, b->stream));
    } else {
        VRS_HIP(ctx, hipGraphLaunch(b->exec, b->stream));
    }
    if (V)
        VRS_HIP(ctx, hipMemcpyAsync(V, b->d_V, sizeof(float2) * b->n_pad_vis, hipMemcpyDefault,
                                    b->stream));
    if (F)
        VRS_HIP(ctx, hipMemcpyAsync(F, b->d_F, sizeof(float) * b->n_ch * g->n_pix, hipMemcpyDefault,
                                    b->stream));
    VRS_HIP(ctx, hipEventRecord(b->ev_out, b->stream));
    VRS_HIP(ctx, hipStreamWaitEvent(ctx->stream, b->ev_out, 0));
    // host destinations (or the pageable kappa source) need the copies finished
    VRS_HIP(ctx, hipStreamSynchronize(b->stream));
    return VRS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- reference compatibility
namespace {
vrs_status io_status(int rc, const char *what, const char *path) {
    switch (rc) {
        case 0: return VRS_OK;
        case -2: return fail(nullptr, VRS_E_ARG, "%s: more rows in %s than the capacity", what, path);
        case -3: return fail(nullptr, VRS_E_IO, "%s: truncated row in %s", what, path);
        default: return fail(nullptr, VRS_E_IO, "%s: cannot open or write %s", what, path);
    }
}
}  // namespace

extern "C" {

vrs_status vrs_shade_box(vrs_ctx *ctx, const float *uvw, int64_t n, int32_t res,
                         float r_moon_ref_km, int32_t *box4) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (!uvw || !box4 || n < 1) return fail(ctx, VRS_E_ARG, "uvw, box4 and n >= 1 are required");
    if (res < 2 || !(r_moon_ref_km > 0)) return fail(ctx, VRS_E_ARG, "res >= 2 and r_moon_ref_km > 0");
    DeviceGuard dg(ctx->device);
    Staged su(ctx, uvw, sizeof(float) * 3 * n, false);
    Staged sb(ctx, box4, sizeof(int32_t) * 4 * n, true);
    if (vrs_status s = su.prepare()) return s;
    if (vrs_status s = sb.prepare()) return s;
    void *partial;
    if (vrs_status s = ws_get(ctx, 3, sizeof(double) * shade_box_partials(), &partial)) return s;
    {
        Timer tm(ctx, "shade_box");
        VRS_HIP(ctx, launch_shade_box((const float *)su.dev, n, res, r_moon_ref_km, (double *)partial,
                                      (int32_t *)sb.dev, ctx->stream));
    }
    return sb.finish();
}

vrs_status vrs_write_uvw_txt(const char *path, const double *uvw, int64_t n, const float *col4) {
    if (!path || (n > 0 && !uvw) || n < 0) return fail(nullptr, VRS_E_ARG, "bad path, uvw or n");
    return io_status(write_uvw_txt(path, uvw, n, col4), "vrs_write_uvw_txt", path);
}

vrs_status vrs_read_uvw_txt(const char *path, int32_t n_cols, double *uvw, float *col4,
                            int64_t cap, int64_t *n_read) {
    if (!path || !uvw || !n_read || cap < 0 || (n_cols != 3 && n_cols != 4))
        return fail(nullptr, VRS_E_ARG, "bad arguments (n_cols must be 3 or 4)");
    *n_read = 0;
    return io_status(read_uvw_txt(path, n_cols, uvw, col4, cap, n_read), "vrs_read_uvw_txt", path);
}

vrs_status vrs_uvw_map_weights(const char *map_path, const double *uvw, int64_t n, float *w,
                               int64_t *n_cells, int64_t *n_skipped, int64_t *n_found) {
    if (!map_path || n < 0 || (n > 0 && (!uvw || !w)))
        return fail(nullptr, VRS_E_ARG, "bad map_path, uvw, w or n");
    switch (uvw_map_weights(map_path, uvw, n, w, n_cells, n_skipped, n_found)) {
        case 0: return VRS_OK;
        case -3: return fail(nullptr, VRS_E_IO, "vrs_uvw_map_weights: no cell count on the first line of %s", map_path);
        case -4: return fail(nullptr, VRS_E_ARG, "vrs_uvw_map_weights: a uvw row is not finite or outside the int32 range");
        case -5: return fail(nullptr, VRS_E_ARG, "vrs_uvw_map_weights: a count <= 0 in %s", map_path);
        default: return fail(nullptr, VRS_E_IO, "vrs_uvw_map_weights: cannot open or read %s", map_path);
    }
}

vrs_status vrs_write_uvw_map_txt(const char *path, const double *uvw, const int32_t *counts,
                                 int64_t n) {
    if (!path || n < 0 || (n > 0 && (!uvw || !counts)))
        return fail(nullptr, VRS_E_ARG, "bad path, uvw, counts or n");
    const int rc = write_uvw_map_txt(path, uvw, counts, n);
    if (rc == -4)
        return fail(nullptr, VRS_E_ARG, "vrs_write_uvw_map_txt: a row outside the int32 range or a count <= 0");
    return io_status(rc, "vrs_write_uvw_map_txt", path);
}

vrs_status vrs_write_shade_txt(const char *path, const int32_t *box4, int64_t n) {
    if (!path || (n > 0 && !box4) || n < 0) return fail(nullptr, VRS_E_ARG, "bad path, box4 or n");
    return io_status(write_shade_txt(path, box4, n), "vrs_write_shade_txt", path);
}

vrs_status vrs_read_shade_txt(const char *path, int32_t *box4, int64_t cap, int64_t *n_read) {
    if (!path || !box4 || !n_read || cap < 0) return fail(nullptr, VRS_E_ARG, "bad arguments");
    *n_read = 0;
    return io_status(read_shade_txt(path, box4, cap, n_read), "vrs_read_shade_txt", path);
}

vrs_status vrs_write_image_txt(const char *path, const float *F, int64_t n) {
    if (!path || (n > 0 && !F) || n < 0) return fail(nullptr, VRS_E_ARG, "bad path, F or n");
    return io_status(write_image_txt(path, F, n), "vrs_write_image_txt", path);
}

vrs_status vrs_read_image_txt(const char *path, float *F, int64_t cap, int64_t *n_read) {
    if (!path || !F || !n_read || cap < 0) return fail(nullptr, VRS_E_ARG, "bad arguments");
    *n_read = 0;
    return io_status(read_image_txt(path, F, cap, n_read), "vrs_read_image_txt", path);
}

vrs_status vrs_write_ex_txt(const char *path, const float *F, int64_t n_pix, const int32_t *nx,
                            int64_t n_cells, int32_t nx_base) {
    if (!path || n_pix < 0 || n_cells < 0 || (n_pix > 0 && !F) || (n_cells > 0 && !nx) ||
        (nx_base != 0 && nx_base != 1))
        return fail(nullptr, VRS_E_ARG, "bad path, F, nx, sizes or nx_base");
    const int rc = write_ex_txt(path, F, n_pix, nx, n_cells, nx_base);
    if (rc == -2) return fail(nullptr, VRS_E_ARG, "vrs_write_ex_txt: a parent pixel is out of range");
    return io_status(rc, "vrs_write_ex_txt", path);
}

vrs_status vrs_write_vis_bin(const char *path, const float *V, int64_t n_maps, int64_t n_vis) {
    if (!path || n_maps < 0 || n_vis < 0 || (n_maps * n_vis > 0 && !V))
        return fail(nullptr, VRS_E_ARG, "bad path, V or sizes");
    return io_status(write_vis_bin(path, V, n_maps, n_vis), "vrs_write_vis_bin", path);
}

vrs_status vrs_read_vis_bin(const char *path, float *V, int64_t cap, int64_t *n_maps,
                            int64_t *n_vis) {
    if (!path || !n_maps || !n_vis || cap < 0) return fail(nullptr, VRS_E_ARG, "bad arguments");
    *n_maps = 0;
    *n_vis = 0;
    const int rc = read_vis_bin(path, V, cap, n_maps, n_vis);
    if (rc == -2)
        return fail(nullptr, VRS_E_ARG, "vrs_read_vis_bin: %lld visibilities exceed cap %lld",
                    (long long)(*n_maps * *n_vis), (long long)cap);
    return io_status(rc, "vrs_read_vis_bin", path);
}

}  // extern "C"

// ---------------------------------------------------------------- uvw gridding
extern "C" vrs_status vrs_grid_uvw(vrs_ctx *ctx, const vrs_geom *g, int32_t flags, int64_t cap,
                                   float *uvw_unique, int32_t *counts, int64_t *n_unique,
                                   float *weight, int32_t *vis_count) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (!n_unique) return fail(ctx, VRS_E_ARG, "n_unique must be non-NULL");
    if (g->from_uvw)
        return fail(ctx, VRS_E_STATE, "gridding needs satellite positions (uvw-row geometry)");
    if ((flags & ~(VRS_GRID_ORDERED | VRS_GRID_LAMBDA_INT)) != 0)
        return fail(ctx, VRS_E_ARG, "unknown grid flags 0x%x", flags);
    if ((uvw_unique || counts) && cap < 1) return fail(ctx, VRS_E_ARG, "cap must be >= 1");
    DeviceGuard dg(ctx->device);
    std::vector<int32_t> pr;
    if (flags & VRS_GRID_ORDERED) {
        for (int m = 0; m < g->n_sat; ++m)
            for (int n = 0; n < g->n_sat; ++n)
                if (m != n) { pr.push_back(m); pr.push_back(n); }
    } else {
        pr = g->pairs;
    }
    const int n_pairs = (int)(pr.size() / 2);
    const int64_t n = (int64_t)n_pairs * g->n_t;
    if (n > ((int64_t)1 << 36)) return fail(ctx, VRS_E_ARG, "too many gridded samples");
    const double lam = (flags & VRS_GRID_LAMBDA_INT) ? (double)(int)(3e8 / g->freq_hz) : g->lambda_m;
    if (!(lam > 0)) return fail(ctx, VRS_E_ARG, "wavelength rounds to zero");
    const double sc = 1e3 / lam;
    // scratch (not the hot path: plain allocations, freed on return)
    struct Buf {
        void *p = nullptr;
        ~Buf() { if (p) (void)hipFree(p); }
    } b_pairs, b_keys, b_tmp, b_rad, b_uniq, b_ukeys, b_cnt, b_vkeys, b_bad;
    VRS_HIP(ctx, hipMalloc(&b_pairs.p, sizeof(int32_t) * pr.size()));
    VRS_HIP(ctx, hipMalloc(&b_keys.p, sizeof(uint64_t) * n));
    VRS_HIP(ctx, hipMalloc(&b_tmp.p, sizeof(uint64_t) * n));
    VRS_HIP(ctx, hipMalloc(&b_rad.p, radix_scratch_bytes(n)));
    VRS_HIP(ctx, hipMalloc(&b_uniq.p, unique_scratch_bytes(n)));
    VRS_HIP(ctx, hipMalloc(&b_bad.p, sizeof(int)));
    VRS_HIP(ctx, hipMemcpyAsync(b_pairs.p, pr.data(), sizeof(int32_t) * pr.size(),
                                hipMemcpyHostToDevice, ctx->stream));
    VRS_HIP(ctx, hipMemsetAsync(b_bad.p, 0, sizeof(int), ctx->stream));
    int64_t nu = 0;
    {
        Timer tm(ctx, "grid");
        VRS_HIP(ctx, grid_build_keys(g->d_pos, g->n_t, (const int2 *)b_pairs.p, n_pairs, sc,
                                     (uint64_t *)b_keys.p, (int *)b_bad.p, ctx->stream));
        VRS_HIP(ctx, radix_sort((uint64_t *)b_keys.p, (uint64_t *)b_tmp.p, n, 0, 8, b_rad.p,
                                ctx->stream, nullptr));
        VRS_HIP(ctx, sorted_unique_count((const uint64_t *)b_keys.p, n, b_uniq.p, &nu, ctx->stream));
    }
    int bad = 0;
    VRS_HIP(ctx, hipMemcpyAsync(&bad, b_bad.p, sizeof bad, hipMemcpyDeviceToHost, ctx->stream));
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (bad) return fail(ctx, VRS_E_ARG, "gridded |uvw| exceeds 2^20 wavelengths");
    *n_unique = nu;
    if (!(uvw_unique || counts || weight || vis_count)) return VRS_OK;
    if ((uvw_unique || counts) && nu > cap)
        return fail(ctx, VRS_E_ARG, "%lld unique cells exceed cap %lld", (long long)nu, (long long)cap);
    VRS_HIP(ctx, hipMalloc(&b_ukeys.p, sizeof(uint64_t) * std::max<int64_t>(nu, 1)));
    VRS_HIP(ctx, hipMalloc(&b_cnt.p, sizeof(int32_t) * std::max<int64_t>(nu, 1)));
    // the sort's second key buffer is free now: it holds the runs' first positions
    VRS_HIP(ctx, sorted_unique_runs((const uint64_t *)b_keys.p, n, b_uniq.p, nu,
                                    (uint64_t *)b_ukeys.p, (int64_t *)b_tmp.p, (int32_t *)b_cnt.p,
                                    ctx->stream));
    if (uvw_unique) {
        Staged su(ctx, uvw_unique, sizeof(float) * 3 * nu, true);
        if (vrs_status s = su.prepare()) return s;
        VRS_HIP(ctx, grid_decode((const uint64_t *)b_ukeys.p, nu, (float *)su.dev, ctx->stream));
        if (vrs_status s = su.finish()) return s;
    }
    if (counts) {
        Staged sc2(ctx, counts, sizeof(int32_t) * nu, true);
        if (vrs_status s = sc2.prepare()) return s;
        VRS_HIP(ctx, hipMemcpyAsync(sc2.dev, b_cnt.p, sizeof(int32_t) * nu,
                                    hipMemcpyDeviceToDevice, ctx->stream));
        if (vrs_status s = sc2.finish()) return s;
    }
    if (weight || vis_count) {
        VRS_HIP(ctx, hipMalloc(&b_vkeys.p, sizeof(uint64_t) * std::max<int64_t>(g->n_vis, 1)));
        VRS_HIP(ctx, grid_build_keys(g->d_pos, g->n_t, g->d_pairs, g->n_base, sc,
                                     (uint64_t *)b_vkeys.p, (int *)b_bad.p, ctx->stream));
        Staged sw(ctx, weight, sizeof(float) * g->n_vis, true);
        Staged svc(ctx, vis_count, sizeof(int32_t) * g->n_vis, true);
        if (vrs_status s = sw.prepare()) return s;
        if (vrs_status s = svc.prepare()) return s;
        VRS_HIP(ctx, grid_vis_weights((const uint64_t *)b_vkeys.p, g->n_vis,
                                      (const uint64_t *)b_ukeys.p, (const int32_t *)b_cnt.p, nu,
                                      (float *)sw.dev, (int32_t *)svc.dev, ctx->stream));
        if (vrs_status s = sw.finish()) return s;
        if (vrs_status s = svc.finish()) return s;
    }
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VRS_OK;
}

// ------------------------------------------------------------ fine-cell sky
// The reference's sub-pixel sky (vissGen10M_shade.cu:121-138, inverse index
// vissGen1M_shade_optim2.cu:59-119): cells with their own direction and FF,
// each belonging to one sky pixel.

vrs_status vrs_geom_fine_sky(vrs_ctx *ctx, vrs_geom *g, int64_t n_cells, const double *lmn,
                             const float *ff, const int32_t *nx, int32_t nx_base) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (n_cells < 1 || n_cells >= ((int64_t)1 << 31))
        return fail(ctx, VRS_E_ARG, "n_cells must be in [1, 2^31)");
    if (!lmn || !ff || !nx) return fail(ctx, VRS_E_ARG, "lmn, ff and nx must be non-NULL");
    if (nx_base != 0 && nx_base != 1) return fail(ctx, VRS_E_ARG, "nx_base must be 0 or 1");
    DeviceGuard dg(ctx->device);
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    g->free_fine();
    Staged sl(ctx, lmn, sizeof(double) * 3 * n_cells, false);
    Staged sf(ctx, ff, sizeof(float) * n_cells, false);
    Staged sn(ctx, nx, sizeof(int32_t) * n_cells, false);
    if (vrs_status s = sl.prepare()) return s;
    if (vrs_status s = sf.prepare()) return s;
    if (vrs_status s = sn.prepare()) return s;
    struct Buf {
        void *p = nullptr;
        ~Buf() { if (p) (void)hipFree(p); }
    } b_keys, b_tmp, b_small;
    const int64_t n = n_cells;
    vrs_status st = VRS_OK;
    auto alloc = [&](void **p, size_t bytes) {
        if (st == VRS_OK && hipMalloc(p, std::max(bytes, (size_t)16)) != hipSuccess) {
            (void)hipGetLastError();
            st = fail(ctx, VRS_E_OOM, "fine-cell sky: device allocation of %zu bytes failed", bytes);
        }
    };
    alloc((void **)&g->d_cell32, sizeof(float4) * n);
    alloc((void **)&g->d_cell64, sizeof(double4) * n);
    alloc((void **)&g->d_fidx, sizeof(int32_t) * n);
    alloc((void **)&g->d_fstart, sizeof(int32_t) * g->n_pad);
    alloc((void **)&g->d_fcount, sizeof(int32_t) * g->n_pad);
    alloc((void **)&g->d_cbar, sizeof(float) * g->n_pad);
    alloc((void **)&g->d_pixcap, sizeof(float2) * g->n_pad);
    alloc((void **)&g->d_fchunk, sizeof(ChunkRec) * g->n_chunk);
    alloc(&b_keys.p, sizeof(uint64_t) * n);
    alloc(&b_tmp.p, sizeof(uint64_t) * n);
    alloc(&b_small.p, radix_scratch_bytes(n) + 256);
    if (st != VRS_OK) {
        g->free_fine();
        return st;
    }
    int *bad = (int *)((char *)b_small.p + radix_scratch_bytes(n));
    hipError_t he = hipMemsetAsync(bad, 0, sizeof(int), ctx->stream);
    if (he == hipSuccess)
        he = fine_build_index((const double *)sl.dev, (const float *)sf.dev, (const int32_t *)sn.dev,
                              nx_base, n, g->n_pix, g->d_ring, g->n_pad, g->d_cell32, g->d_cell64,
                              (uint64_t *)b_keys.p, (uint64_t *)b_tmp.p, b_small.p, bad,
                              g->d_fidx, g->d_fstart, g->d_fcount, ctx->stream);
    int hb = 0;
    if (he == hipSuccess) he = hipMemcpyAsync(&hb, bad, sizeof hb, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) {
        g->free_fine();
        return fail(ctx, VRS_E_HIP, "fine-cell sky index: %s", hipGetErrorString(he));
    }
    if (hb) {
        g->free_fine();
        return fail(ctx, VRS_E_ARG, "fine-cell sky: %s",
                    (hb & 1)   ? "a parent pixel index is out of range"
                    : (hb & 2) ? "a sky pixel has no cells"
                               : "a cell direction is not a unit vector");
    }
    g->n_cells = n;
    GeomDev gd = g->dev();
    gd.fine = {g->d_fstart, g->d_fcount, g->d_fidx, g->d_cell32, g->d_cell64, nullptr};
    he = launch_fine_pixels(gd, g->d_cbar, g->d_pixcap, g->d_fchunk, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) {
        g->free_fine();
        return fail(ctx, VRS_E_HIP, "fine-cell sky bounds: %s", hipGetErrorString(he));
    }
    return VRS_OK;
}

vrs_status vrs_fine_index(vrs_ctx *ctx, const vrs_geom *g, int32_t *count_loc, int32_t *nxq,
                          float *pixel_mean) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (g->n_cells == 0) return fail(ctx, VRS_E_STATE, "no fine-cell sky attached to the geometry");
    DeviceGuard dg(ctx->device);
    Staged sc(ctx, count_loc, sizeof(int32_t) * g->n_pix, true);
    Staged sq(ctx, nxq, sizeof(int32_t) * g->n_cells, true);
    Staged sm(ctx, pixel_mean, sizeof(float) * g->n_pix, true);
    if (vrs_status s = sc.prepare()) return s;
    if (vrs_status s = sq.prepare()) return s;
    if (vrs_status s = sm.prepare()) return s;
    if (count_loc)
        VRS_HIP(ctx, fine_count_ring(g->d_fcount, g->d_ring, g->n_pad, (int32_t *)sc.dev, ctx->stream));
    if (nxq)
        VRS_HIP(ctx, hipMemcpyAsync(sq.dev, g->d_fidx, sizeof(int32_t) * g->n_cells,
                                    hipMemcpyDeviceToDevice, ctx->stream));
    if (pixel_mean) {
        // chunk order -> RING order through the same gather the recon output uses
        const GeomDev gd = g->dev();
        VRS_HIP(ctx, launch_from_chunk_order(gd, g->n_pix, g->d_cbar, (float *)sm.dev, ctx->stream));
    }
    if (vrs_status s = sc.finish()) return s;
    if (vrs_status s = sq.finish()) return s;
    if (vrs_status s = sm.finish()) return s;
    VRS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VRS_OK;
}

vrs_status vrs_visgen_fine(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, float *V) {
    if (vrs_status s = check_ctx(ctx)) return s;
    if (vrs_status s = check_geom(ctx, g)) return s;
    if (vrs_status s = check_knobs(ctx, k)) return s;
    if (!V) return fail(ctx, VRS_E_ARG, "V must be non-NULL");
    if (g->n_cells == 0) return fail(ctx, VRS_E_STATE, "no fine-cell sky attached to the geometry");
    if (k->precision == VRS_MFMA_SPLIT16)
        return fail(ctx, VRS_E_ARG, "the fine-cell sky is one map: use plain or mixed precision");
    DeviceGuard dg(ctx->device);
    Staged sV(ctx, V, sizeof(float) * 2 * g->n_vis, true);
    if (vrs_status s = sV.prepare()) return s;
    GeomDev gd = g->dev();
    gd.chunk = g->d_fchunk;
    gd.fine = {g->d_fstart, g->d_fcount, g->d_fidx, g->d_cell32, g->d_cell64, g->d_pixcap};
    const bool f64 = k->precision == VRS_MIXED_F64PHASE;
    const int vt = visgen_tile_vis();
    const int64_t n_tiles = (g->n_vis + vt - 1) / vt;
    const int grid = persistent_grid(ctx, visgen_fine_kernel_ptr(f64), n_tiles);
    {
        Timer tm(ctx, "visgen_fine");
        VRS_HIP(ctx, launch_visgen_fine(gd, phase_knobs(k), f64, g->d_cbar, (float2 *)sV.dev, grid,
                                        ctx->stream));
    }
    return sV.finish();
}
