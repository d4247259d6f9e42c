// rf_runtime.cpp -- host side of the C ABI (include/rayfactor.h): contexts,
// device scenes, the dispatch that replaces runCL_GPURNG's OpenCL launches
// (RayFactorCL/main.cpp:54-251), batched persistent solves, the autotuner
// and the streaming service.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "rayfactor.h"
#include "rf_internal.hpp"

using namespace rf;
using rfd::FastPrim;
using rfd::EmitFrame;
using rfd::Emitter;

// ---------------------------------------------------------------------------
// Context and scene objects
// ---------------------------------------------------------------------------
static constexpr int kFan = 4;          // lanes of concurrent per-emitter launches
static constexpr size_t kRing = 4096;   // queue counters per lane

struct rf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    char name[256] = {0};
    int n_cu = 0;
    int clock_mhz = 0;
    int tuned_rpt = 0, tuned_block = 0, tuned_transform = 0;
    int tuned_tile = 0;                  // band records in LDS: 0 untuned, 1 off, 2 on
    // scratch device buffers
    GeomDesc *d_geoms = nullptr; size_t geoms_cap = 0;
    WorkUnit *d_units = nullptr; size_t units_cap = 0;
    // queue counters: one ring of kRing per fan lane (a lane reuses its own
    // slots only, after its earlier launches in stream order)
    unsigned *d_queue = nullptr;
    size_t queue_next[kFan] = {};
    // Fan-out of independent per-emitter launches over kFan lanes: lane 0 is
    // `stream`, the others extra streams (default priority).  They fork from
    // `stream` and join back into it before anything reads results.
    hipStream_t fan[kFan] = {};
    hipEvent_t fan_ev[kFan] = {};
    hipEvent_t fork_ev = nullptr;
    bool fan_open = false;
    int fan_next = 0;
    unsigned long long *d_stats = nullptr;
    // Persistent solves repeated with the same scenes and options replay a
    // HIP graph of [zero fills, queue kernel]: the first call runs eagerly, the
    // second captures, later ones replay.  d_geoms / d_units hold that
    // solve's descriptors and units while graph_key is set.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;     // solve captured in graph_exec
    std::vector<uint64_t> last_key;      // last persistent solve run eagerly
    int64_t graph_replays = 0;
    int fail_in = 0;                     // test hook: the fail_in-th next solve fails (0 = off)
    LaunchCfg last_cfg{};                // shape of the last solve launch (reported in rf_timing)
    void *scratch_ptr = nullptr;         // this device's allocation-free scratch (see scratch_setup)
    // live objects owned by the context (released by rf_ctx_destroy)
    std::set<rf_scene *> scenes;
    std::set<rf_stream *> streams;
};

// A band stack of the fast scan, for cost estimates.
struct StackInfo {
    int32_t bands;
    bool sorted;                         // z offsets sorted: searched by default
};

struct rf_scene {
    rf_ctx *ctx = nullptr;
    uint64_t serial = 0;                 // unique per scene object (graph keys)
    int n = 0, n_bvs = 0;
    std::vector<rf_primitive> prims;
    std::vector<rf_bounding_volume> bvs;
    rf_primitive *d_prims = nullptr;
    FastPrim *d_fast = nullptr, *d_fast_orig = nullptr, *d_fast_bvs = nullptr;
    rf_bounding_volume *d_bvs = nullptr;
    BvMeta *d_bv_meta = nullptr;
    unsigned long long *d_hits = nullptr;
    float *d_quads = nullptr;            // MFMA form: primitive quads
    int32_t *d_pos2orig = nullptr, *d_orig2pos = nullptr;
    FastPrim *d_bvc_fast = nullptr;      // binned BV form: children per BV, type-segmented
    int32_t *d_bvc_seg = nullptr;
    CylRun *d_cyl_runs = nullptr;        // fast form: cylinder runs (band stacks)
    BandRec *d_bands = nullptr;
    Emitter *d_emitters = nullptr;       // primitives with their emission frames
    BvhNode *d_bvh_nodes = nullptr;      // fast form: hierarchy over the per-test primitives
    FastPrim *d_bvh_prims = nullptr;     // their leaves' records
    GeomDesc *d_desc = nullptr;          // single-geometry descriptor
    GeomDesc desc{};                     // host copy
    int64_t flops_per_ray = 0;           // sum over primitives of F_type (plain kernels)
    int64_t flops_per_ray_fast = 0;      // the fast VALU scan's (world z-plane rings at kFlopsZRing)
    int64_t flops_per_ray_nonstack = 0;  // the fast scan's without the band stacks' cylinders (counted on the device)
    int64_t flops_per_ray_cnonstack = 0; // the same for the compat scan's stacks (caller order)
    CylRun *d_c_runs = nullptr;          // compat scan: runs over the caller's order
    BandRec *d_c_bands = nullptr;
    int32_t *d_c_run_of = nullptr;       // run index of each caller position (BV child scans)
    int64_t flops_bvs = 0;               // sum over bounding surfaces of F_type (BV tests per ray)
    CylRun *d_bvc_runs = nullptr;        // binned BV form: cylinder runs of each surface's children
    int32_t *d_bvc_run_off = nullptr;
    BandRec *d_bvc_bands = nullptr;
    uint32_t *d_bv_cflops = nullptr;
    double flops_per_test_avg = 0.0;     // BV accounting
    bool has_search = false;             // a fast-form band stack with sorted offsets
    bool bvh_on = false;                 // fast form: a hierarchy over the per-test primitives
    int bvh_depth = 0;                   // its depth: traversal stack entries a walk needs
    std::vector<StackInfo> stack_info;   // the fast scan's band stacks
    int32_t stack_prims = 0;             // primitives inside them
    int32_t n_nonstack = 0;              // primitives outside the fast scan's band stacks
    int32_t n_cnonstack = 0;             // the same for the compat / FP64 scans (caller order)
};

// Binned BV form: at most this many bounding surfaces per scene (LDS buckets;
// the kernel packs the BV index into 16 bits).
static constexpr int kMaxBvBins = 512;

static void drop_graph(rf_ctx *ctx);
static int fan_join(rf_ctx *ctx);

static thread_local std::string g_err;

static int set_err(rf_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg; else g_err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err(ctx, RF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int32_t rf_abi_version(void) { return RF_ABI_VERSION; }

extern "C" const char *rf_status_string(int32_t s)
{
    switch (s) {
    case RF_OK: return "ok";
    case RF_ERR_INVALID_ARG: return "invalid argument";
    case RF_ERR_INVALID_SCENE: return "invalid scene";
    case RF_ERR_HIP: return "HIP runtime error";
    case RF_ERR_OUT_OF_MEMORY: return "out of device memory";
    case RF_ERR_OVERFLOW: return "integer overflow";
    case RF_ERR_IO: return "I/O error";
    case RF_ERR_PARSE: return "scene parse error";
    case RF_ERR_CAPACITY: return "capacity exceeded";
    case RF_ERR_NO_DEVICE: return "no such HIP device";
    default: return "unknown status";
    }
}

extern "C" void rf_solve_opts_default(rf_solve_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->math = RF_MATH_COMPAT;
    o->rng = RF_RNG_REFERENCE;
    o->use_bvs = 0;
    o->budget = RF_BUDGET_REFERENCE;
    o->n_samples = 1000000;
    o->form = RF_FORM_PERSISTENT;
    o->refine = 1;
    o->count_tests = 1;
}

extern "C" const char *rf_last_error(const rf_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

// ---------------------------------------------------------------------------
// Scratch of the allocation-free solve (rf_solve_noalloc): its queue counter
// and statistics live in this fixed device array, one copy per GPU, held until
// the process ends.  Host code uses the address hipGetSymbolAddress returns
// for the current device (scratch_ptr), never the array's name.
// Layout: [0, 4) queue counter; [8, 8 + 8 kNStats) statistics.
// ---------------------------------------------------------------------------
static constexpr size_t kScratchBytes = 256;
static constexpr size_t kScratchStatsOff = 8;
static_assert(kScratchStatsOff + kNStats * sizeof(unsigned long long) <= kScratchBytes, "scratch layout");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch, looked up once per device with the
// device current (called by rf_ctx_create).
static hipError_t scratch_setup(int dev, void **ptr)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

extern "C" int32_t rf_ctx_create(int32_t dev, rf_ctx **out)
{
    if (!out) return set_err(nullptr, RF_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return set_err(nullptr, RF_ERR_NO_DEVICE, "no HIP device visible");
    if (dev < 0 || dev >= count) return set_err(nullptr, RF_ERR_NO_DEVICE, "device index out of range");
    rf_ctx *c = new rf_ctx;
    c->device = dev;
    hipError_t e = hipSetDevice(dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
    for (int i = 1; i < kFan && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&c->fan[i], hipStreamNonBlocking);   // default priority
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fan_ev[i], hipEventDisableTiming);
    }
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_stats, kNStats * sizeof(unsigned long long));
    if (e == hipSuccess) e = scratch_setup(dev, &c->scratch_ptr);
    if (e != hipSuccess) {
        g_err = std::string("context creation failed: ") + hipGetErrorString(e);
        delete c;
        return RF_ERR_HIP;
    }
    std::snprintf(c->name, sizeof c->name, "%s", prop.gcnArchName);
    c->n_cu = prop.multiProcessorCount;
    c->clock_mhz = prop.clockRate / 1000;
    *out = c;
    return RF_OK;
}

extern "C" int32_t rf_ctx_destroy(rf_ctx *c)
{
    if (!c) return RF_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) {
        (void)fan_join(c);
        (void)hipStreamSynchronize(c->stream);
    }
    // streams own scenes of their own; release them first, then remaining scenes
    while (!c->streams.empty()) rf_stream_destroy(*c->streams.begin());
    while (!c->scenes.empty()) rf_scene_destroy(*c->scenes.begin());
    drop_graph(c);
    (void)hipFree(c->d_geoms);
    (void)hipFree(c->d_units);
    (void)hipFree(c->d_queue);
    (void)hipFree(c->d_stats);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    for (int i = 1; i < kFan; ++i) {
        if (c->fan_ev[i]) (void)hipEventDestroy(c->fan_ev[i]);
        if (c->fan[i]) (void)hipStreamDestroy(c->fan[i]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RF_OK;
}

extern "C" int32_t rf_device_info(const rf_ctx *c, char *name, size_t len, int32_t *n_cu, int32_t *clock_mhz)
{
    if (!c) return RF_ERR_INVALID_ARG;
    if (name && len) std::snprintf(name, len, "%s", c->name);
    if (n_cu) *n_cu = c->n_cu;
    if (clock_mhz) *clock_mhz = c->clock_mhz;
    return RF_OK;
}

extern "C" int32_t rf_selftest_math(rf_ctx *ctx, uint64_t mismatches[4], uint32_t first[4])
{
    if (!ctx || !mismatches || !first) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = fan_join(ctx);
    if (rc) return rc;
    unsigned long long *d = nullptr;
    HIPCHK(ctx, hipMalloc((void **)&d, 4 * sizeof(unsigned long long) + 4 * sizeof(unsigned)));
    unsigned *d_first = reinterpret_cast<unsigned *>(d + 4);
    hipError_t e = launch_fill_u64(d, 4, ctx->stream);
    if (e == hipSuccess) e = launch_fill_u32(d_first, 0xffffffffu, 4, ctx->stream);
    if (e == hipSuccess) e = launch_selftest_math(d, d_first, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(mismatches, d, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(first, d_first, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(ctx, RF_ERR_HIP, std::string("selftest: ") + hipGetErrorString(e));
    return RF_OK;
}

extern "C" int32_t rf_selftest_sincos(rf_ctx *ctx, double max_err[2])
{
    if (!ctx || !max_err) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = fan_join(ctx);
    if (rc) return rc;
    unsigned *d = nullptr;
    HIPCHK(ctx, hipMalloc((void **)&d, 2 * sizeof(unsigned)));
    unsigned h[2] = {0u, 0u};
    hipError_t e = launch_fill_u32(d, 0u, 2, ctx->stream);
    if (e == hipSuccess) e = launch_selftest_sincos(d, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(ctx, RF_ERR_HIP, std::string("selftest: ") + hipGetErrorString(e));
    for (int k = 0; k < 2; ++k) {
        float f;
        std::memcpy(&f, &h[k], sizeof f);
        max_err[k] = f;
    }
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Scene upload
// ---------------------------------------------------------------------------
static bool valid_type(int32_t t) { return t >= RF_CYLINDER && t < RF_NONE; }

static FastPrim make_fast(const float invm[12], int32_t type, float u1, float u2, int32_t orig)
{
    FastPrim f;
    std::memcpy(f.r, invm, sizeof f.r);
    f.k0 = 0.f; f.k1 = 0.f; f.orig = orig; f.type = type;
    switch (type) {
    case RF_CYLINDER:
        f.r[11] = invm[11] - 0.5f;                       // accept |z - 1/2| <= 1/2
        break;
    case RF_FRUSTUM: {
        const float k = u1 - 1.0f;                       // w = 1 + (u1-1) z: cone radius
        for (int j = 0; j < 4; ++j) f.r[8 + j] = k * invm[8 + j];
        f.r[11] += 1.0f;
        f.k0 = 0.5f * (1.0f + u1);
        f.k1 = 0.5f * std::fabs(u1 - 1.0f);
        break;
    }
    case RF_DISC: f.k0 = 0.f; f.k1 = 1.f; break;
    case RF_ANNULUS: f.k0 = u1; f.k1 = u2; break;
    default: break;
    }
    return f;
}

// A disc or annulus whose plane is z = z0 in world coordinates and whose
// object x-y is a similarity (scale s, rotation or reflection) of world x-y:
// the fast scan tests it in world coordinates (zring_loop).  The record is
// re-encoded: r[0] = z0, r[1], r[2] = the centre's world x, y, k0, k1 = the
// r^2 range in world units (the object range / s^2).  Checked in FP64 on the
// FP32 matrix: the z row must not depend on world x or y beyond the round-off
// of a rotation by a multiple of 90 degrees (|m20|, |m21| at most 1e-7 |m22|;
// float sin(pi) is -8.7e-8), and the x-y block's columns must have equal norms
// and be orthogonal to 1e-6 relative.  The tilt so ignored moves the plane by
// at most 1e-7 |x| in object z, the order of the FP32 rounding of the
// object-frame test.
static bool zring_record(const float invm[12], int32_t type, float u1, float u2, int32_t orig, FastPrim &f)
{
    if (type != RF_DISC && type != RF_ANNULUS) return false;
    const double a = invm[0], b = invm[1], c = invm[4], d = invm[5];
    const double m22 = invm[10];
    if (!(std::fabs(m22) > 0.0) || !std::isfinite(m22)) return false;
    if (std::fabs((double)invm[8]) > 1e-7 * std::fabs(m22) || std::fabs((double)invm[9]) > 1e-7 * std::fabs(m22))
        return false;
    const double s2 = a * a + c * c, s2b = b * b + d * d;
    if (!(s2 > 0.0) || !std::isfinite(s2)) return false;
    if (std::fabs(s2b - s2) > 1e-6 * s2 || std::fabs(a * b + c * d) > 1e-6 * s2) return false;
    const double z0 = -(double)invm[11] / m22;
    // object x, y on the plane: a X + b Y + e0, c X + d Y + e1 (the world-z
    // terms are constants there); the centre solves them = 0
    const double e0 = (double)invm[2] * z0 + (double)invm[3], e1 = (double)invm[6] * z0 + (double)invm[7];
    const double det = a * d - b * c;
    if (!(std::fabs(det) > 0.0)) return false;
    const double x0 = (-e0 * d + b * e1) / det, y0 = (-a * e1 + c * e0) / det;
    const double k0 = type == RF_DISC ? 0.0 : (double)u1, k1 = type == RF_DISC ? 1.0 : (double)u2;
    f = FastPrim{};
    f.r[0] = (float)z0; f.r[1] = (float)x0; f.r[2] = (float)y0;
    f.k0 = (float)(k0 / s2); f.k1 = (float)(k1 / s2);
    f.orig = orig; f.type = type;
    return std::isfinite(f.r[0]) && std::isfinite(f.r[1]) && std::isfinite(f.r[2]) && std::isfinite(f.k0) &&
           std::isfinite(f.k1);
}

static int seg_of(int32_t type)
{
    switch (type) {
    case RF_CYLINDER: return SEG_CYL;
    case RF_FRUSTUM: return SEG_FRUSTUM;
    case RF_SPHERE: return SEG_SPHERE;
    case RF_RECTANGLE: return SEG_RECT;
    default: return SEG_RING;
    }
}

// Output rows (4 x [x, y, z, 1] coefficients) of one primitive in the MFMA
// form; see the kernel's quad layout.  pad: a slot that never hits.
static void quad_rows(const rf_primitive *p, float rows[4][4])
{
    std::memset(rows, 0, sizeof(float) * 16);
    if (!p) {
        rows[0][3] = 1e3f;    // x = 1000, zero direction: quadrics see A = 0 (no root)
        rows[2][3] = 1e3f;    // z = 1000, zero direction: planes see t = -inf
        return;
    }
    const float *iv = p->invm;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) rows[r][k] = iv[4 * r + k];
    switch (p->type) {
    case RF_CYLINDER:
        rows[2][3] = iv[11] - 0.5f;
        break;
    case RF_FRUSTUM: {
        const float k = p->util1 - 1.0f;
        for (int c = 0; c < 4; ++c) { rows[3][c] = iv[8 + c]; rows[2][c] = k * iv[8 + c]; }
        rows[2][3] += 1.0f;                         // w = 1 + (u1-1) z
        rows[3][3] = iv[11] - 0.5f;                 // z - 1/2
        break;
    }
    case RF_ANNULUS: {
        const float rs = 1.0f / std::sqrt(p->util2);  // outer radius -> 1
        for (int c = 0; c < 4; ++c) { rows[0][c] *= rs; rows[1][c] *= rs; }
        rows[3][3] = p->util1 / p->util2;           // inner radius^2, normalised
        break;
    }
    default:
        break;                                      // disc (r^2 in [0,1]), sphere, rectangle
    }
}

// Cylinder runs over records [cb, ce) of a type-segmented array: maximal
// groups of consecutive cylinders whose folded rows other than the z offset
// r[11] are bitwise equal; two or more form a band stack, consecutive
// singletons one per-test run.  Returns minus the per-test flops per ray of
// the stacked cylinders (the kernel counts a stack's flops on the device).
static int64_t build_cyl_runs(const std::vector<FastPrim> &recs, int32_t cb, int32_t ce, std::vector<CylRun> &runs);

// The stack run [p, p + n) whose band offsets are off[0..n): stack = 2 when
// they are finite and non-decreasing, 3 when finite and non-increasing (the
// kernels may then search the bands), 1 otherwise (every band is walked);
// inv = (n - 1) / (off[n-1] - off[0]), the bands per unit offset of an evenly
// spaced stack (0 when the offsets are equal).
template <class Off>
static CylRun stack_run(int32_t p, int32_t n, Off off)
{
    bool up = true, down = true;
    for (int32_t k = 0; k < n; ++k) {
        if (!std::isfinite(off(k))) return CylRun{p, p + n, 1, 0.f};
        if (k > 0) {
            up = up && off(k - 1) <= off(k);
            down = down && off(k - 1) >= off(k);
        }
    }
    if (!up && !down) return CylRun{p, p + n, 1, 0.f};
    const double span = (double)off(n - 1) - (double)off(0);
    const float inv = span != 0.0 ? (float)((double)(n - 1) / span) : 0.f;
    CylRun r{p, p + n, up ? 2 : 3, std::isfinite(inv) ? inv : 0.f};
    r.r0 = off(0);
    return r;
}

static uint32_t float_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, sizeof u);
    return u;
}

// Analytic band positions of a searched stack (CylRun, fast form): the model
// progression runs from the first offset with the end-to-end mean spacing,
// raised to |delta| = 1 when rounding left it just below (bands that abut);
// an_margin is the largest deviation of an offset from the model plus the
// FP64 evaluation's rounding and the FP32 threshold compares' (both far below
// the FP32 per-band tests' own rounding, which the kernel adds per ray).
// Stacks whose bands overlap (|delta| < 1), are far apart (|delta| > 64) or
// deviate by more than 1/32 of a band keep the searched records.
template <class Off, class Orig>
static void analytic_positions(CylRun &r, bool axial, Off off, Orig orig)
{
    const int32_t n = r.end - r.begin;
    if (r.stack < 2 || !axial || n < 2) return;
    double delta = ((double)off(n - 1) - (double)off(0)) / (double)(n - 1);
    const double ad = std::fabs(delta);
    if (!(ad >= 1.0 - 1e-4) || ad > 64.0) return;
    if (ad < 1.0) delta = delta < 0 ? -1.0 : 1.0;
    double dev = 0;
    for (int32_t k = 0; k < n; ++k)
        dev = std::max(dev, std::fabs((double)off(k) - ((double)off(0) + (double)k * delta)));
    const double margin = dev + 1e-9 * (1.0 + (double)n * std::fabs(delta)) + 4e-7;
    if (!(margin < 1.0 / 32)) return;
    const double ninv = -1.0 / delta;
    r.an_ninv = ninv;
    r.an_ainv = (float)std::fabs(ninv);
    r.an_margin = (float)(margin * (1.0 + 1e-6));
    r.orig0 = orig(0);
    int32_t step = n >= 2 ? orig(1) - orig(0) : 0;
    if (step != 1 && step != -1) step = 0;
    for (int32_t k = 1; k < n && step; ++k)
        if (orig(k) != r.orig0 + k * step) step = 0;
    r.ostep = step;
}

// The same over the caller's order for the compat scan: a stack is two or
// more consecutive cylinders with bitwise equal invm[0..10]; everything else
// forms per-test runs.  Returns minus the per-test flops of the stacked cylinders.
static int64_t build_compat_runs(const rf_primitive *prims, int32_t n, std::vector<CylRun> &runs)
{
    int64_t delta = 0;
    auto stack_end = [&](int32_t p) {
        int32_t q = p + 1;
        if (prims[p].type != RF_CYLINDER) return q;
        while (q < n && prims[q].type == RF_CYLINDER &&
               std::memcmp(prims[q].invm, prims[p].invm, 11 * sizeof(float)) == 0)
            ++q;
        return q;
    };
    int32_t p = 0;
    while (p < n) {
        int32_t q = stack_end(p);
        if (q - p >= 2) {
            const auto off = [&](int32_t k) { return prims[p + k].invm[11]; };
            CylRun r = stack_run(p, q - p, off);
            const bool axial = ((float_bits(prims[p].invm[8]) | float_bits(prims[p].invm[9])) << 1) == 0u;
            analytic_positions(r, axial, off, [&](int32_t k) { return p + k; });
            runs.push_back(r);
            delta -= (q - p) * flops_per_test(RF_CYLINDER);
            p = q;
            continue;
        }
        const int32_t s0 = p;
        while (p < n && (q = stack_end(p)) - p < 2) p = q;
        runs.push_back(CylRun{s0, p, 0, 0});
    }
    return delta;
}

static int64_t build_cyl_runs(const std::vector<FastPrim> &recs, int32_t cb, int32_t ce, std::vector<CylRun> &runs)
{
    int64_t delta = 0;
    auto stack_end = [&](int32_t p) {
        int32_t q = p + 1;
        while (q < ce && std::memcmp(recs[q].r, recs[p].r, 11 * sizeof(float)) == 0) ++q;
        return q;
    };
    int32_t p = cb;
    while (p < ce) {
        int32_t q = stack_end(p);
        if (q - p >= 2) {
            const auto off = [&](int32_t k) { return recs[p + k].r[11]; };
            CylRun r = stack_run(p, q - p, off);
            // the stack_loop test of an axial z row: integer tests on the bits (+-0)
            const bool axial = ((float_bits(recs[p].r[8]) | float_bits(recs[p].r[9])) << 1) == 0u;
            analytic_positions(r, axial, off, [&](int32_t k) { return recs[p + k].orig; });
            runs.push_back(r);
            delta -= (q - p) * flops_per_test(RF_CYLINDER);
            p = q;
            continue;
        }
        const int32_t s0 = p;
        while (p < ce && (q = stack_end(p)) - p < 2) p = q;
        runs.push_back(CylRun{s0, p, 0, 0});
    }
    return delta;
}

// ---------------------------------------------------------------------------
// Hierarchies of the fast form's per-test primitives (BvhNode)
// ---------------------------------------------------------------------------
// Scenes with at least this many primitives outside band stacks get them.
static constexpr int32_t kBvhMin = 48;
// Primitives per leaf: on 300- to 2000-primitive soups one primitive per
// leaf traced 1.2x the rays per second of two, 1.4x of four, 2x of eight.
static constexpr int kBvhLeaf = 1;

struct WBox {
    double lo[3], hi[3];
};

static double box_area(const WBox &b)
{
    const double x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return x * y + y * z + z * x;
}

// Centroid bins per axis of the surface-area split.
static constexpr int kSahBins = 32;

// World box of a primitive: its object-space shape's box (the unit cylinder
// and the frustum of radii 1 and |util1| over z in [0, 1], the unit sphere,
// the disc, annulus and rectangle in z = 0) through the inverse of invm (the
// map every hit test applies), in FP64.  False when invm is singular or the
// box is not finite.
static bool prim_box(const rf_primitive &p, WBox &b)
{
    double lo[3] = {-1, -1, 0}, hi[3] = {1, 1, 1};
    switch (p.type) {
    case RF_FRUSTUM: {
        const double rm = std::max(1.0, std::fabs((double)p.util1));
        lo[0] = lo[1] = -rm; hi[0] = hi[1] = rm;
        break;
    }
    case RF_SPHERE: lo[2] = -1; break;
    case RF_DISC:
    case RF_RECTANGLE: hi[2] = 0; break;
    case RF_ANNULUS: {
        const double ro = std::sqrt(std::max({(double)p.util1, (double)p.util2, 0.0}));
        lo[0] = lo[1] = -ro; hi[0] = hi[1] = ro; hi[2] = 0;
        break;
    }
    default: break;
    }
    // object = A world + v; world = A^-1 (object - v) by the adjugate
    double A[3][3], v[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) A[r][c] = p.invm[4 * r + c];
        v[r] = p.invm[4 * r + 3];
    }
    double adj[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
            adj[r][c] = A[r1][c1] * A[r2][c2] - A[r1][c2] * A[r2][c1];
        }
    const double det = A[0][0] * adj[0][0] + A[0][1] * adj[1][0] + A[0][2] * adj[2][0];
    if (!std::isfinite(det) || std::fabs(det) < 1e-300) return false;
    for (int a = 0; a < 3; ++a) { b.lo[a] = INFINITY; b.hi[a] = -INFINITY; }
    for (int c = 0; c < 8; ++c) {
        const double o[3] = {((c & 1) ? hi[0] : lo[0]) - v[0], ((c & 2) ? hi[1] : lo[1]) - v[1],
                             ((c & 4) ? hi[2] : lo[2]) - v[2]};
        for (int a = 0; a < 3; ++a) {
            const double w = (adj[a][0] * o[0] + adj[a][1] * o[1] + adj[a][2] * o[2]) / det;
            b.lo[a] = std::min(b.lo[a], w);
            b.hi[a] = std::max(b.hi[a], w);
        }
    }
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(b.lo[a]) || !std::isfinite(b.hi[a])) return false;
    return true;
}

// Box frame of a scene's hierarchy nodes: world x = c + s h for a stored
// half-precision h, with s a power of two.
struct BoxFrame {
    float c[3];
    float s;
};

// IEEE half-precision bit pattern of the largest half <= v (up = false) or
// the smallest half >= v (up = true), for |v| <= 2^15.
static uint16_t half_bits(double v, bool up)
{
    if (v == 0.0) return 0;
    int e = 0;
    (void)std::frexp(std::fabs(v), &e);          // |v| in [2^(e-1), 2^e)
    int ex = std::max(e - 1, -14);                // binade; below 2^-14 the subnormal spacing
    double m = std::ldexp(v, 10 - ex);            // exact: in quanta of the binade
    m = up ? std::ceil(m) : std::floor(m);
    const uint16_t sign = m < 0 ? 0x8000 : 0;
    double a = std::fabs(m);
    if (a == 0.0) return sign;
    if (a >= 2048.0) { a *= 0.5; ++ex; }          // rounded up into the next binade
    if (a < 1024.0) return (uint16_t)(sign | (uint16_t)a);   // subnormal
    return (uint16_t)(sign | (uint16_t)((ex + 15) << 10) | (uint16_t)((int)a - 1024));
}

// Value of a half-precision bit pattern (finite).
static double half_value(uint16_t h)
{
    const int ex = (h >> 10) & 31, mt = h & 1023;
    const double v = ex ? std::ldexp(1024.0 + mt, ex - 25) : std::ldexp((double)mt, -24);
    return (h & 0x8000) ? -v : v;
}

// Hierarchy over items (traversal positions with their padded boxes and type
// segments), split by the surface-area heuristic (the split of the centroid
// bins that minimises count x surface area summed over both sides) while the
// depth allows, appended to nodes / pos.  A leaf holds at most
// kBvhLeaf positions of one type segment; a group small enough for a leaf
// but of mixed segments is split between segments.  Returns the root node
// index and the depth (the most stack entries a traversal needs).
static int32_t build_bvh(std::vector<int32_t> items, const std::vector<WBox> &box, const std::vector<int> &kind,
                         const BoxFrame &fr, std::vector<BvhNode> &nodes, std::vector<int32_t> &pos, int &depth)
{
    struct Ref {
        int32_t c, n;
        WBox b;
    };
    const auto bound = [&](const int32_t *it, size_t cnt) {
        WBox u;
        for (int a = 0; a < 3; ++a) { u.lo[a] = INFINITY; u.hi[a] = -INFINITY; }
        for (size_t k = 0; k < cnt; ++k)
            for (int a = 0; a < 3; ++a) {
                u.lo[a] = std::min(u.lo[a], box[it[k]].lo[a]);
                u.hi[a] = std::max(u.hi[a], box[it[k]].hi[a]);
            }
        return u;
    };
    // a child's box in the frame, corners rounded outward; its reference word
    const auto put_box = [&](BvhNode &nd, int k, const WBox &b) {
        for (int a = 0; a < 3; ++a) {
            nd.box[6 * k + a] = half_bits((b.lo[a] - fr.c[a]) / fr.s, false);
            nd.box[6 * k + 3 + a] = half_bits((b.hi[a] - fr.c[a]) / fr.s, true);
        }
    };
    const auto ref_of = [](const Ref &x) {
        if (x.c < 0) return BVH_NO_CHILD;
        const uint32_t cnt = (uint32_t)x.n & 255u, seg = (uint32_t)x.n >> 8;
        return ((uint32_t)x.c & BVH_NO_CHILD) | cnt << BVH_INDEX_BITS | seg << (BVH_INDEX_BITS + 3);
    };
    const auto node_of = [&](int32_t self, const Ref &l, const Ref &r) {
        BvhNode &nd = nodes[self];
        put_box(nd, 0, l.b);
        put_box(nd, 1, r.b);
        nd.ref[0] = ref_of(l);
        nd.ref[1] = ref_of(r);
        WBox u;
        for (int a = 0; a < 3; ++a) { u.lo[a] = std::min(l.b.lo[a], r.b.lo[a]); u.hi[a] = std::max(l.b.hi[a], r.b.hi[a]); }
        return Ref{self, 0, u};
    };
    depth = 0;
    std::function<Ref(int32_t *, size_t, int)> rec = [&](int32_t *it, size_t cnt, int level) -> Ref {
        depth = std::max(depth, level);
        bool one = true;
        for (size_t k = 1; k < cnt; ++k) one = one && kind[it[k]] == kind[it[0]];
        if (cnt <= (size_t)kBvhLeaf && one) {
            const int32_t first = (int32_t)pos.size();
            pos.insert(pos.end(), it, it + cnt);
            return Ref{first, (int32_t)cnt | (kind[it[0]] << 8), bound(it, cnt)};
        }
        size_t mid = cnt / 2;
        if (cnt <= (size_t)kBvhLeaf) {
            // mixed segments: split at the segment boundary nearest the middle
            std::sort(it, it + cnt, [&](int32_t u, int32_t v) { return kind[u] < kind[v] || (kind[u] == kind[v] && u < v); });
            size_t best = 0;
            for (size_t k = 1; k < cnt; ++k)
                if (kind[it[k]] != kind[it[k - 1]] &&
                    (best == 0 || std::llabs((long long)(2 * k) - (long long)cnt) < std::llabs((long long)(2 * best) - (long long)cnt)))
                    best = k;
            mid = best;
        } else {
            WBox cb;
            for (int a = 0; a < 3; ++a) { cb.lo[a] = INFINITY; cb.hi[a] = -INFINITY; }
            for (size_t k = 0; k < cnt; ++k)
                for (int a = 0; a < 3; ++a) {
                    const double m = 0.5 * (box[it[k]].lo[a] + box[it[k]].hi[a]);
                    cb.lo[a] = std::min(cb.lo[a], m);
                    cb.hi[a] = std::max(cb.hi[a], m);
                }
            // surface-area split over kSahBins centroid bins per axis, while
            // the levels left allow an unbalanced split; else the median on
            // the longest centroid axis
            int lg = 0;
            while (((size_t)1 << lg) < cnt) ++lg;
            bool split = false;
            if (level + lg + 4 < BVH_MAX_DEPTH) {
                double best = INFINITY;
                int bax = -1, bs = 0;
                for (int a = 0; a < 3; ++a) {
                    const double ext = cb.hi[a] - cb.lo[a];
                    if (!(ext > 0.0)) continue;
                    int64_t n_in[kSahBins] = {};
                    WBox bb[kSahBins];
                    for (WBox &w : bb)
                        for (int d = 0; d < 3; ++d) { w.lo[d] = INFINITY; w.hi[d] = -INFINITY; }
                    for (size_t k = 0; k < cnt; ++k) {
                        const WBox &w = box[it[k]];
                        const int b = std::min(kSahBins - 1, (int)(kSahBins * (0.5 * (w.lo[a] + w.hi[a]) - cb.lo[a]) / ext));
                        ++n_in[b];
                        for (int d = 0; d < 3; ++d) { bb[b].lo[d] = std::min(bb[b].lo[d], w.lo[d]); bb[b].hi[d] = std::max(bb[b].hi[d], w.hi[d]); }
                    }
                    double ra[kSahBins];
                    int64_t rn[kSahBins];
                    WBox acc;
                    for (int d = 0; d < 3; ++d) { acc.lo[d] = INFINITY; acc.hi[d] = -INFINITY; }
                    int64_t num = 0;
                    for (int b = kSahBins - 1; b >= 1; --b) {
                        num += n_in[b];
                        for (int d = 0; d < 3; ++d) { acc.lo[d] = std::min(acc.lo[d], bb[b].lo[d]); acc.hi[d] = std::max(acc.hi[d], bb[b].hi[d]); }
                        ra[b] = num ? box_area(acc) : 0.0;
                        rn[b] = num;
                    }
                    for (int d = 0; d < 3; ++d) { acc.lo[d] = INFINITY; acc.hi[d] = -INFINITY; }
                    num = 0;
                    for (int b = 0; b + 1 < kSahBins; ++b) {
                        num += n_in[b];
                        for (int d = 0; d < 3; ++d) { acc.lo[d] = std::min(acc.lo[d], bb[b].lo[d]); acc.hi[d] = std::max(acc.hi[d], bb[b].hi[d]); }
                        if (!num || !rn[b + 1]) continue;
                        const double c = (double)num * box_area(acc) + (double)rn[b + 1] * ra[b + 1];
                        if (c < best) { best = c; bax = a; bs = b + 1; }
                    }
                }
                if (bax >= 0) {
                    const double ext = cb.hi[bax] - cb.lo[bax];
                    int32_t *m = std::partition(it, it + cnt, [&](int32_t u) {
                        const WBox &w = box[u];
                        return std::min(kSahBins - 1, (int)(kSahBins * (0.5 * (w.lo[bax] + w.hi[bax]) - cb.lo[bax]) / ext)) < bs;
                    });
                    mid = (size_t)(m - it);
                    split = mid > 0 && mid < cnt;
                }
            }
            if (!split) {
                mid = cnt / 2;
                int ax = 0;
                for (int a = 1; a < 3; ++a)
                    if (cb.hi[a] - cb.lo[a] > cb.hi[ax] - cb.lo[ax]) ax = a;
                std::nth_element(it, it + mid, it + cnt, [&](int32_t u, int32_t v) {
                    const double cu = box[u].lo[ax] + box[u].hi[ax], cv = box[v].lo[ax] + box[v].hi[ax];
                    return cu < cv || (cu == cv && u < v);
                });
            }
        }
        const int32_t self = (int32_t)nodes.size();
        nodes.push_back(BvhNode{});
        const Ref l = rec(it, mid, level + 1), r = rec(it + mid, cnt - mid, level + 1);
        return node_of(self, l, r);
    };
    if (items.empty()) return -1;
    bool one = true;
    for (size_t k = 1; k < items.size(); ++k) one = one && kind[items[k]] == kind[items[0]];
    if (items.size() <= (size_t)kBvhLeaf && one) {
        // a root whose one child is the leaf
        const int32_t self = (int32_t)nodes.size();
        nodes.push_back(BvhNode{});
        const Ref l = rec(items.data(), items.size(), 1);
        BvhNode &nd = nodes[self];
        put_box(nd, 0, l.b);
        for (int a = 0; a < 3; ++a) { nd.box[6 + a] = 0x3C00; nd.box[9 + a] = 0xBC00; }   // +1, -1: empty
        nd.ref[0] = ref_of(l);
        nd.ref[1] = BVH_NO_CHILD;
        return self;
    }
    return rec(items.data(), items.size(), 0).c;
}

// Emission frame of a primitive (EmitFrame): kind 1 for a cylinder whose
// invm rows A, B are orthogonal, of equal length and orthogonal to its axis a
// (m's third column), to a relative 1e-6; kind 0 (the per-ray frame) for every
// other primitive.  Built in FP64 from the FP32 records.
static EmitFrame emit_frame(const rf_primitive &p)
{
    EmitFrame f{};
    if (p.type != RF_CYLINDER) return f;
    const double A[3] = {p.invm[0], p.invm[1], p.invm[2]}, B[3] = {p.invm[4], p.invm[5], p.invm[6]};
    const double a[3] = {p.m[2], p.m[6], p.m[10]};
    const auto dot = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double la = std::sqrt(dot(A, A)), lb = std::sqrt(dot(B, B)), lx = std::sqrt(dot(a, a));
    if (!(la > 0 && lb > 0 && lx > 0) || !std::isfinite(la * lb * lx)) return f;
    const double tol = 1e-6;
    if (std::fabs(la - lb) > tol * la || std::fabs(dot(A, B)) > tol * la * lb || std::fabs(dot(a, A)) > tol * lx * la ||
        std::fabs(dot(a, B)) > tol * lx * lb)
        return f;
    if (p.util2 == 0.0f || !std::isfinite(p.util2)) return f;
    const double sg = p.util2 > 0 ? 1.0 : -1.0;
    double na[3], nb[3], ax[3];
    for (int j = 0; j < 3; ++j) { na[j] = sg * A[j] / la; nb[j] = sg * B[j] / lb; ax[j] = a[j] / lx; }
    const auto cross = [](const double *x, const double *y, double *z) {
        z[0] = x[1] * y[2] - x[2] * y[1];
        z[1] = x[2] * y[0] - x[0] * y[2];
        z[2] = x[0] * y[1] - x[1] * y[0];
    };
    double pa[3], pb[3];
    cross(na, ax, pa);
    cross(nb, ax, pb);
    f.kind = 1;
    for (int j = 0; j < 3; ++j) {
        f.ax[j] = (float)ax[j]; f.na[j] = (float)na[j]; f.nb[j] = (float)nb[j];
        f.pa[j] = (float)pa[j]; f.pb[j] = (float)pb[j];
    }
    return f;
}

extern "C" int64_t rf_scene_footprint(int32_t n, int32_t nbv)
{
    if (n < 0 || nbv < 0) return -1;
    return (int64_t)n * (sizeof(rf_primitive) + sizeof(Emitter) + 2 * sizeof(FastPrim) + 64 + 8 + 2 * sizeof(BandRec) +
                         sizeof(BvhNode) + sizeof(FastPrim) +
                         2 * sizeof(CylRun) + 4) + 9 * 256 +
           (int64_t)nbv * (sizeof(rf_bounding_volume) + sizeof(FastPrim) + sizeof(BvMeta) + (N_SEG + 1) * 4 +
                           4 + 8) +
           (nbv ? (int64_t)n * (sizeof(FastPrim) + sizeof(BandRec) + sizeof(CylRun)) + 6 * 256 : 0) +
           (int64_t)n * n * 8 + (int64_t)sizeof(GeomDesc);
}

static void free_scene_mem(rf_scene *s)
{
    (void)hipFree(s->d_quads); (void)hipFree(s->d_pos2orig); (void)hipFree(s->d_orig2pos);
    (void)hipFree(s->d_bvc_fast); (void)hipFree(s->d_bvc_seg);
    (void)hipFree(s->d_cyl_runs); (void)hipFree(s->d_bands); (void)hipFree(s->d_emitters);
    (void)hipFree(s->d_bvh_nodes); (void)hipFree(s->d_bvh_prims);
    (void)hipFree(s->d_c_runs); (void)hipFree(s->d_c_bands); (void)hipFree(s->d_c_run_of);
    (void)hipFree(s->d_bvc_runs); (void)hipFree(s->d_bvc_run_off); (void)hipFree(s->d_bvc_bands);
    (void)hipFree(s->d_bv_cflops);
    (void)hipFree(s->d_prims); (void)hipFree(s->d_fast); (void)hipFree(s->d_fast_orig);
    (void)hipFree(s->d_fast_bvs); (void)hipFree(s->d_bvs); (void)hipFree(s->d_bv_meta);
    (void)hipFree(s->d_hits); (void)hipFree(s->d_desc);
}

static int validate_scene(const rf_primitive *prims, int32_t n, const rf_bounding_volume *bvs, int32_t nbv,
                          std::string &why)
{
    if (n <= 0 || !prims) { why = "scene has no primitives"; return RF_ERR_INVALID_SCENE; }
    if (n > (1 << 20)) { why = "more than 2^20 primitives"; return RF_ERR_CAPACITY; }
    if (nbv < 0 || (nbv > 0 && !bvs)) { why = "bad bounding-volume array"; return RF_ERR_INVALID_ARG; }
    for (int32_t i = 0; i < n; ++i)
        if (!valid_type(prims[i].type)) { why = "primitive " + std::to_string(i) + " has an unknown type"; return RF_ERR_INVALID_SCENE; }
    for (int32_t i = 0; i < nbv; ++i) {
        if (!valid_type(bvs[i].type)) { why = "bounding volume " + std::to_string(i) + " has an unknown type"; return RF_ERR_INVALID_SCENE; }
        if (bvs[i].objStartIndex < 0 || bvs[i].objEndIndex > n || bvs[i].objStartIndex > bvs[i].objEndIndex) {
            why = "bounding volume " + std::to_string(i) + " has an invalid child range";
            return RF_ERR_INVALID_SCENE;
        }
    }
    return RF_OK;
}

extern "C" int32_t rf_scene_create(rf_ctx *ctx, const rf_primitive *prims, int32_t n, const rf_bounding_volume *bvs,
                                   int32_t nbv, rf_scene **out)
{
    if (!ctx || !out) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL context or output");
    *out = nullptr;
    std::string why;
    int rc = validate_scene(prims, n, bvs, nbv, why);
    if (rc) return set_err(ctx, rc, why);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    static std::atomic<uint64_t> next_serial{1};
    rf_scene *s = new rf_scene;
    s->serial = next_serial.fetch_add(1);
    s->ctx = ctx; s->n = n; s->n_bvs = nbv;
    s->prims.assign(prims, prims + n);
    if (nbv) s->bvs.assign(bvs, bvs + nbv);

    // fast form: a hierarchy over the per-test primitives when many lie
    // outside band stacks.  Their scenes group the cylinder segment: every
    // set of two or more cylinders with bitwise equal rows (all but the z
    // offset) contiguous, in caller order, sets in order of their first
    // member, then the cylinders of unique rows; so the band stacks come
    // first and exact ties between a band and a per-test cylinder go to the
    // band in every scan.  Their rings all keep the object-frame test.
    bool bvh_on = false;
    std::vector<int32_t> cyl_order;   // bvh_on: the cylinder segment's caller indices
    {
        std::vector<int32_t> ci;
        for (int32_t i = 0; i < n; ++i)
            if (prims[i].type == RF_CYLINDER) ci.push_back(i);
        const auto row = [&](int32_t i) {
            std::array<uint32_t, 11> k;
            const FastPrim f = make_fast(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i);
            std::memcpy(k.data(), f.r, sizeof k);
            return k;
        };
        std::map<std::array<uint32_t, 11>, std::vector<int32_t>> sets;
        std::vector<std::array<uint32_t, 11>> first;
        for (int32_t i : ci) {
            auto &v = sets[row(i)];
            if (v.empty()) first.push_back(row(i));
            v.push_back(i);
        }
        int32_t grouped = 0;
        for (const auto &k : first) {
            const auto &v = sets[k];
            if (v.size() >= 2) { cyl_order.insert(cyl_order.end(), v.begin(), v.end()); grouped += (int32_t)v.size(); }
        }
        for (const auto &k : first) {
            const auto &v = sets[k];
            if (v.size() == 1) cyl_order.push_back(v[0]);
        }
        bvh_on = n - grouped >= kBvhMin;
        if (!bvh_on) cyl_order.clear();
    }
    // type-segmented traversal order (stable within a type)
    std::vector<FastPrim> fast, fast_orig(n), fast_bvs(nbv);
    std::vector<BvMeta> meta(nbv);
    int32_t seg_begin[N_SEG], seg_end[N_SEG], zring_end = 0, n_zring = 0;
    for (int k = 0; k < N_SEG; ++k) {
        seg_begin[k] = (int32_t)fast.size();
        if (k == SEG_RING) {
            // world z-plane rings first (zring_loop), then the others
            for (int32_t i = 0; i < n; ++i) {
                FastPrim z;
                if (!bvh_on && seg_of(prims[i].type) == k &&
                    zring_record(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i, z))
                    fast.push_back(z);
            }
            zring_end = (int32_t)fast.size();
            n_zring = zring_end - seg_begin[k];
            for (int32_t i = 0; i < n; ++i) {
                FastPrim z;
                if (seg_of(prims[i].type) == k &&
                    (bvh_on || !zring_record(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i, z)))
                    fast.push_back(make_fast(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i));
            }
        } else if (k == SEG_CYL && bvh_on) {
            for (int32_t i : cyl_order)
                fast.push_back(make_fast(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i));
        } else {
            for (int32_t i = 0; i < n; ++i)
                if (seg_of(prims[i].type) == k)
                    fast.push_back(make_fast(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i));
        }
        seg_end[k] = (int32_t)fast.size();
    }
    int64_t fsum = 0;
    for (int32_t i = 0; i < n; ++i) {
        fast_orig[i] = make_fast(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i);
        fsum += flops_per_test(prims[i].type);
    }
    int64_t fbv = 0;
    for (int32_t i = 0; i < nbv; ++i) {
        fast_bvs[i] = make_fast(bvs[i].invm, bvs[i].type, bvs[i].util1, bvs[i].util2, i);
        meta[i] = BvMeta{bvs[i].settings, bvs[i].objStartIndex, bvs[i].objEndIndex, bvs[i].type};
        fbv += flops_per_test(bvs[i].type);
    }
    // MFMA form: quads of 4 primitives per type segment, padded
    std::vector<float> quads;
    std::vector<int32_t> pos2orig, orig2pos(n, -1);
    int32_t qseg_begin[N_SEG], qseg_end[N_SEG];
    for (int k = 0; k < N_SEG; ++k) {
        qseg_begin[k] = (int32_t)(pos2orig.size() / 4);
        std::vector<int32_t> members;
        for (int32_t i = 0; i < n; ++i)
            if (seg_of(prims[i].type) == k) members.push_back(i);
        while (members.size() % 4) members.push_back(-1);
        for (size_t q = 0; q < members.size(); q += 4) {
            float a[64];
            for (int slot = 0; slot < 4; ++slot) {
                const int32_t i = members[q + slot];
                float rows[4][4];
                quad_rows(i >= 0 ? &prims[i] : nullptr, rows);
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) a[16 * c + 4 * slot + r] = rows[r][c];
                if (i >= 0) orig2pos[i] = (int32_t)pos2orig.size();
                pos2orig.push_back(i);
            }
            quads.insert(quads.end(), a, a + 64);
        }
        qseg_end[k] = (int32_t)(pos2orig.size() / 4);
    }
    // binned BV form: the children of each BV in type-segmented order; usable
    // when no BV asks for the intermediate check (runBVGPU's settings == 1)
    std::vector<FastPrim> bvc;
    std::vector<int32_t> bvc_seg;
    bool binned = nbv >= 1 && nbv <= kMaxBvBins;
    for (int32_t b = 0; b < nbv && binned; ++b)
        if (bvs[b].settings == 1) binned = false;
    if (binned) {
        bvc_seg.resize((size_t)nbv * (N_SEG + 1));
        for (int32_t b = 0; b < nbv; ++b) {
            for (int k = 0; k < N_SEG; ++k) {
                bvc_seg[(size_t)b * (N_SEG + 1) + k] = (int32_t)bvc.size();
                for (int32_t i = bvs[b].objStartIndex; i < bvs[b].objEndIndex; ++i)
                    if (seg_of(prims[i].type) == k)
                        bvc.push_back(make_fast(prims[i].invm, prims[i].type, prims[i].util1, prims[i].util2, i));
            }
            bvc_seg[(size_t)b * (N_SEG + 1) + N_SEG] = (int32_t)bvc.size();
        }
    }
    // band stacks of the plain scan; band records by traversal position after
    // one pad record (the kernel's prefetch reads the record before a stack)
    std::vector<CylRun> runs;
    const int64_t fnonstack = fsum + build_cyl_runs(fast, seg_begin[SEG_CYL], seg_end[SEG_CYL], runs);
    for (const CylRun &r : runs) {
        s->has_search = s->has_search || r.stack >= 2;
        if (r.stack) {
            s->stack_info.push_back(StackInfo{r.end - r.begin, r.stack >= 2});
            s->stack_prims += r.end - r.begin;
        }
    }
    std::vector<BandRec> bands(n + 1, BandRec{0.f, 0});
    for (int32_t k = seg_begin[SEG_CYL]; k < seg_end[SEG_CYL]; ++k) bands[k + 1] = BandRec{fast[k].r[11], fast[k].orig};
    // hierarchy over the per-test positions (the cylinders of per-test runs,
    // every other segment whole), leaves of one segment, boxes padded by
    // 1e-4 of the scene's extent plus 1e-5 of its largest coordinate: far
    // beyond the FP32 error of a hit point (FP32 test or FP64 re-solve, a few
    // ulps of the coordinates) and of the kernel's slab entry and exit
    // distances (one FMA of a box face with the reciprocal direction, a few
    // ulps of the largest coordinate over |c|), so a pruned box holds no hit
    // at or before the best one; the nodes store the padded boxes in half
    // precision relative to the scene's centre, rounded outward
    std::vector<BvhNode> bvh_nodes;
    std::vector<int32_t> bvh_pos;
    std::vector<FastPrim> bvh_prims;
    int32_t bvh_root = -1;
    int bvh_depth = 0;
    BoxFrame frame{{0.0f, 0.0f, 0.0f}, 1.0f};
    if (bvh_on) {
        std::vector<WBox> box(n);
        WBox all;
        for (int a = 0; a < 3; ++a) { all.lo[a] = INFINITY; all.hi[a] = -INFINITY; }
        for (int32_t p = 0; p < n && bvh_on; ++p) {
            bvh_on = prim_box(prims[fast[p].orig], box[p]);
            for (int a = 0; a < 3 && bvh_on; ++a) {
                all.lo[a] = std::min(all.lo[a], box[p].lo[a]);
                all.hi[a] = std::max(all.hi[a], box[p].hi[a]);
            }
        }
        if (bvh_on) {
            double diag = 0, big = 0;
            for (int a = 0; a < 3; ++a) {
                diag += (all.hi[a] - all.lo[a]) * (all.hi[a] - all.lo[a]);
                big = std::max({big, std::fabs(all.lo[a]), std::fabs(all.hi[a])});
            }
            const double pad = 1e-4 * std::sqrt(diag) + 1e-5 * big + 1e-30;
            for (WBox &b : box)
                for (int a = 0; a < 3; ++a) { b.lo[a] -= pad; b.hi[a] += pad; }
            std::vector<int32_t> items;
            std::vector<int> kind(n, 0);
            for (int k = 0; k < N_SEG; ++k) {
                for (int32_t p = seg_begin[k]; p < seg_end[k]; ++p) kind[p] = k;
                if (k == SEG_CYL) {
                    for (const CylRun &r : runs)
                        if (!r.stack)
                            for (int32_t p = r.begin; p < r.end; ++p) items.push_back(p);
                } else {
                    for (int32_t p = seg_begin[k]; p < seg_end[k]; ++p) items.push_back(p);
                }
            }
            // the nodes' box frame: the centre of the scene's box, and a
            // power-of-two scale that puts every padded corner within 2^14
            // of it in half precision
            double reach = 0.0;
            for (int a = 0; a < 3; ++a) {
                frame.c[a] = (float)(0.5 * (all.lo[a] + all.hi[a]));
                reach = std::max({reach, std::fabs(all.lo[a] - pad - frame.c[a]), std::fabs(all.hi[a] + pad - frame.c[a])});
            }
            int se = 0;
            (void)std::frexp(reach / 16384.0, &se);
            frame.s = std::ldexp(1.0f, se);
            int deepest = 0;
            bvh_root = build_bvh(items, box, kind, frame, bvh_nodes, bvh_pos, deepest);
            // a traversal keeps at most one entry per level on its stack;
            // child references hold 25-bit indices
            if (deepest >= BVH_MAX_DEPTH || bvh_nodes.size() >= BVH_NO_CHILD / 2 || bvh_pos.size() >= BVH_NO_CHILD / 2)
                bvh_on = false;
            // every stored box holds its padded box (outward rounding)
            for (size_t q = 0; q < bvh_nodes.size() && bvh_on; ++q)
                for (int k = 0; k < 2; ++k)
                    for (int a = 0; a < 3; ++a)
                        if (!std::isfinite(half_value(bvh_nodes[q].box[6 * k + a])) ||
                            !std::isfinite(half_value(bvh_nodes[q].box[6 * k + 3 + a])))
                            bvh_on = false;
            bvh_depth = deepest;
            // leaf records in leaf order, each carrying its traversal position
            for (int32_t p : bvh_pos) {
                FastPrim f = fast[p];
                f.type = p;
                bvh_prims.push_back(f);
            }
        }
        if (!bvh_on) {
            bvh_nodes.clear();
            bvh_prims.clear();
            bvh_root = -1;
        }
    }
    s->bvh_on = bvh_on;
    s->bvh_depth = bvh_on ? bvh_depth : 0;
    // compat scan: stacks over the caller's order, band records {invm[11], index}
    std::vector<CylRun> c_runs;
    const int64_t fcnonstack = fsum + build_compat_runs(prims, n, c_runs);
    std::vector<BandRec> c_bands(n + 1, BandRec{0.f, 0});
    for (int32_t k = 0; k < n; ++k) c_bands[k + 1] = BandRec{prims[k].invm[11], k};
    // the same for the children of each bounding surface (binned BV form), and
    // per surface: the children's flops per ray (per test; with stacks, those
    // outside them), the children, and the children outside the stacks
    std::vector<CylRun> bvc_runs;
    std::vector<int32_t> bvc_run_off;
    std::vector<BandRec> bvc_bands;
    std::vector<uint32_t> bv_cflops;
    if (binned) {
        bvc_run_off.resize(nbv + 1);
        bv_cflops.resize(4 * (size_t)nbv);
        bvc_bands.assign(bvc.size() + 1, BandRec{0.f, 0});
        for (int32_t b = 0; b < nbv; ++b) {
            const int32_t *sg = &bvc_seg[(size_t)b * (N_SEG + 1)];
            bvc_run_off[b] = (int32_t)bvc_runs.size();
            const int64_t d = build_cyl_runs(bvc, sg[SEG_CYL], sg[SEG_CYL + 1], bvc_runs);
            int64_t f = 0;
            for (int32_t k = sg[0]; k < sg[N_SEG]; ++k) f += flops_per_test(bvc[k].type);
            bv_cflops[4 * b] = (uint32_t)f;
            bv_cflops[4 * b + 1] = (uint32_t)(f + d);
            bv_cflops[4 * b + 2] = (uint32_t)(sg[N_SEG] - sg[0]);
            bv_cflops[4 * b + 3] = (uint32_t)(sg[N_SEG] - sg[0] + d / flops_per_test(RF_CYLINDER));
            for (int32_t k = sg[SEG_CYL]; k < sg[SEG_CYL + 1]; ++k) bvc_bands[k + 1] = BandRec{bvc[k].r[11], bvc[k].orig};
        }
        bvc_run_off[nbv] = (int32_t)bvc_runs.size();
    }
    // the fast VALU scan tests world z-plane rings with kFlopsZRing flops each
    // (and one reciprocal per ray) instead of the reference formula's
    const int64_t fzring = n_zring ? kFlopsZRingRay - (int64_t)n_zring * (flops_per_test(RF_DISC) - kFlopsZRing) : 0;
    s->flops_per_ray = fsum;
    s->flops_per_ray_fast = fsum + fzring;
    s->flops_per_ray_nonstack = fnonstack + fzring;
    s->flops_per_ray_cnonstack = fcnonstack;
    s->n_nonstack = n - (int32_t)((fsum - fnonstack) / flops_per_test(RF_CYLINDER));
    s->n_cnonstack = n - (int32_t)((fsum - fcnonstack) / flops_per_test(RF_CYLINDER));
    s->flops_per_test_avg = (double)(fsum + fbv) / (double)(n + nbv);
    s->flops_bvs = fbv;

    hipError_t e = hipSuccess;
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e != hipSuccess || bytes == 0) return;
        e = hipMalloc(dst, bytes);
        if (e == hipSuccess) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    up((void **)&s->d_prims, prims, sizeof(rf_primitive) * n);
    up((void **)&s->d_fast, fast.data(), sizeof(FastPrim) * n);
    up((void **)&s->d_fast_orig, fast_orig.data(), sizeof(FastPrim) * n);
    if (nbv) {
        up((void **)&s->d_bvs, bvs, sizeof(rf_bounding_volume) * nbv);
        up((void **)&s->d_fast_bvs, fast_bvs.data(), sizeof(FastPrim) * nbv);
        up((void **)&s->d_bv_meta, meta.data(), sizeof(BvMeta) * nbv);
    }
    up((void **)&s->d_quads, quads.data(), quads.size() * sizeof(float));
    up((void **)&s->d_pos2orig, pos2orig.data(), pos2orig.size() * sizeof(int32_t));
    up((void **)&s->d_orig2pos, orig2pos.data(), orig2pos.size() * sizeof(int32_t));
    up((void **)&s->d_cyl_runs, runs.data(), runs.size() * sizeof(CylRun));
    up((void **)&s->d_c_runs, c_runs.data(), c_runs.size() * sizeof(CylRun));
    up((void **)&s->d_c_bands, c_bands.data(), c_bands.size() * sizeof(BandRec));
    std::vector<int32_t> c_run_of(n > 0 ? n : 1, 0);
    for (size_t j = 0; j < c_runs.size(); ++j)
        for (int32_t k = c_runs[j].begin; k < c_runs[j].end; ++k) c_run_of[k] = (int32_t)j;
    up((void **)&s->d_c_run_of, c_run_of.data(), c_run_of.size() * sizeof(int32_t));
    up((void **)&s->d_bands, bands.data(), bands.size() * sizeof(BandRec));
    up((void **)&s->d_bvh_nodes, bvh_nodes.data(), bvh_nodes.size() * sizeof(BvhNode));
    up((void **)&s->d_bvh_prims, bvh_prims.data(), bvh_prims.size() * sizeof(FastPrim));
    std::vector<Emitter> emitters(n);
    for (int32_t i = 0; i < n; ++i) {
        std::memcpy(static_cast<rf_primitive *>(&emitters[i]), &prims[i], sizeof(rf_primitive));
        emitters[i].f = emit_frame(prims[i]);
    }
    up((void **)&s->d_emitters, emitters.data(), emitters.size() * sizeof(Emitter));
    if (binned) {
        up((void **)&s->d_bvc_runs, bvc_runs.data(), bvc_runs.size() * sizeof(CylRun));
        up((void **)&s->d_bvc_run_off, bvc_run_off.data(), bvc_run_off.size() * sizeof(int32_t));
        up((void **)&s->d_bvc_bands, bvc_bands.data(), bvc_bands.size() * sizeof(BandRec));
        up((void **)&s->d_bv_cflops, bv_cflops.data(), bv_cflops.size() * sizeof(uint32_t));
        up((void **)&s->d_bvc_fast, bvc.data(), bvc.size() * sizeof(FastPrim));
        up((void **)&s->d_bvc_seg, bvc_seg.data(), bvc_seg.size() * sizeof(int32_t));
    }
    if (e == hipSuccess) e = hipMalloc((void **)&s->d_hits, (size_t)n * n * sizeof(unsigned long long));
    if (e == hipSuccess) e = launch_fill_u64(s->d_hits, (size_t)n * n, ctx->stream);
    GeomDesc &d = s->desc;
    d.prims = s->d_prims; d.fast = s->d_fast; d.fast_orig = s->d_fast_orig;
    d.bvs = s->d_bvs; d.fast_bvs = s->d_fast_bvs; d.bv_meta = s->d_bv_meta;
    d.hits = s->d_hits; d.n = n; d.n_bvs = nbv; d.geometry_id = 0; d.zring_end = zring_end;
    for (int k = 0; k < N_SEG; ++k) { d.seg_begin[k] = seg_begin[k]; d.seg_end[k] = seg_end[k]; }
    d.quads = s->d_quads; d.pos2orig = s->d_pos2orig; d.orig2pos = s->d_orig2pos;
    for (int k = 0; k < N_SEG; ++k) { d.qseg_begin[k] = qseg_begin[k]; d.qseg_end[k] = qseg_end[k]; }
    d.bvc_fast = s->d_bvc_fast; d.bvc_seg = s->d_bvc_seg; d.bv_binned = binned ? 1 : 0;
    d.n_cyl_runs = (int32_t)runs.size(); d.cyl_runs = s->d_cyl_runs;
    d.bands = s->d_bands ? s->d_bands + 1 : nullptr;
    d.n_c_runs = (int32_t)c_runs.size(); d.c_runs = s->d_c_runs;
    d.c_bands = s->d_c_bands ? s->d_c_bands + 1 : nullptr;
    d.c_run_of = s->d_c_run_of;
    d.bvc_runs = s->d_bvc_runs; d.bvc_run_off = s->d_bvc_run_off;
    d.bvc_bands = s->d_bvc_bands ? s->d_bvc_bands + 1 : nullptr;
    d.bv_cflops = s->d_bv_cflops;
    d.emitters = s->d_emitters;
    d.bvh_nodes = s->d_bvh_nodes;
    d.bvh_prims = s->d_bvh_prims;
    d.bvh_root = bvh_root;
    d.bvh_on = bvh_on ? (s->stack_prims == 0 ? 2 : 1) : 0;
    for (int a = 0; a < 3; ++a) d.bvh_c[a] = frame.c[a];
    d.bvh_s = frame.s;
    up((void **)&s->d_desc, &d, sizeof d);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? RF_ERR_OUT_OF_MEMORY : RF_ERR_HIP;
        set_err(ctx, code, std::string("scene upload failed: ") + hipGetErrorString(e));
        free_scene_mem(s);
        delete s;
        return code;
    }
    ctx->scenes.insert(s);
    *out = s;
    return RF_OK;
}

extern "C" int32_t rf_scene_destroy(rf_scene *s)
{
    if (!s) return RF_ERR_INVALID_ARG;
    (void)hipSetDevice(s->ctx->device);
    (void)fan_join(s->ctx);                      // launches on side lanes may still read the scene
    (void)hipStreamSynchronize(s->ctx->stream);
    s->ctx->scenes.erase(s);
    free_scene_mem(s);
    delete s;
    return RF_OK;
}

extern "C" int32_t rf_scene_size(const rf_scene *s, int32_t *n, int32_t *nbv)
{
    if (!s) return RF_ERR_INVALID_ARG;
    if (n) *n = s->n;
    if (nbv) *nbv = s->n_bvs;
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Launch geometry (main.cpp:101,172-173)
// ---------------------------------------------------------------------------
extern "C" int32_t rf_nthreads(int32_t ray_density, int32_t n_samples, int32_t count, int32_t wgsize, int64_t *out)
{
    if (!out || count <= 0 || wgsize <= 0) return RF_ERR_INVALID_ARG;
    const int64_t prod = (int64_t)ray_density * (int64_t)n_samples;
    if (prod > INT32_MAX || prod < INT32_MIN) return RF_ERR_OVERFLOW;
    const int64_t nt = prod / count;          // int arithmetic in the reference
    if (nt < 0) return RF_ERR_OVERFLOW;       // would wrap through size_t
    const int64_t rem = nt % wgsize;
    *out = rem > wgsize / 2 ? nt + wgsize - rem : nt - rem;
    return RF_OK;
}

static int check_opts(rf_ctx *ctx, const rf_solve_opts *o)
{
    if (!o) return set_err(ctx, RF_ERR_INVALID_ARG, "opts is NULL");
    if (o->math < RF_MATH_COMPAT || o->math > RF_MATH_FAST) return set_err(ctx, RF_ERR_INVALID_ARG, "bad math mode");
    if (o->rng != RF_RNG_REFERENCE && o->rng != RF_RNG_SEEDED) return set_err(ctx, RF_ERR_INVALID_ARG, "bad rng mode");
    if (o->budget != RF_BUDGET_REFERENCE && o->budget != RF_BUDGET_EXPLICIT) return set_err(ctx, RF_ERR_INVALID_ARG, "bad budget mode");
    if (o->form < RF_FORM_PER_EMITTER || o->form > RF_FORM_REFERENCE) return set_err(ctx, RF_ERR_INVALID_ARG, "bad form");
    if (o->budget == RF_BUDGET_EXPLICIT && o->rays_per_emitter < 0) return set_err(ctx, RF_ERR_INVALID_ARG, "negative ray budget");
    if (o->rays_per_thread && o->rays_per_thread != 1 && o->rays_per_thread != 2 && o->rays_per_thread != 4 &&
        o->rays_per_thread != 8)
        return set_err(ctx, RF_ERR_INVALID_ARG, "rays_per_thread must be 0, 1, 2, 4 or 8");
    if (o->transform < RF_TRANSFORM_AUTO || o->transform > RF_TRANSFORM_MFMA)
        return set_err(ctx, RF_ERR_INVALID_ARG, "bad transform");
    if (!(o->refine_margin >= 0.0f && o->refine_margin <= 0.25f))
        return set_err(ctx, RF_ERR_INVALID_ARG, "refine_margin must lie in [0, 0.25]");
    if (o->per_test < 0 || o->per_test > 2) return set_err(ctx, RF_ERR_INVALID_ARG, "per_test must be 0, 1 or 2");
    if (o->block_size && (o->block_size < 64 || o->block_size > 256 || o->block_size % 64))
        return set_err(ctx, RF_ERR_INVALID_ARG, "block_size must be 0 or a multiple of 64 in [64, 256]");
    if (o->count_tests != 0 && o->count_tests != 1) return set_err(ctx, RF_ERR_INVALID_ARG, "count_tests must be 0 or 1");
    if (o->band_tile < -1) return set_err(ctx, RF_ERR_INVALID_ARG, "band_tile must be -1, 0 or positive");
    if (o->hierarchy != 0 && o->hierarchy != -1) return set_err(ctx, RF_ERR_INVALID_ARG, "hierarchy must be 0 or -1");
    return RF_OK;
}

// Work items per emitter of a scene under the options' budget.
static int emitter_budget(rf_ctx *ctx, const rf_scene *s, const rf_solve_opts *o, std::vector<int64_t> &n_wi,
                          int64_t rays_override = -1)
{
    n_wi.assign(s->n, 0);
    if (o->budget == RF_BUDGET_REFERENCE) {
        // every emitter uses objects[i].rayDensity with i = the device index 0 (main.cpp:172)
        int64_t nt;
        int rc = rf_nthreads(s->prims[0].rayDensity, o->n_samples, 4, 128, &nt);
        if (rc) return set_err(ctx, rc, "rayDensity * n_samples overflows int32 (main.cpp:172)");
        if (nt > INT32_MAX) return set_err(ctx, RF_ERR_OVERFLOW, "work items exceed the reference's int gid");
        std::fill(n_wi.begin(), n_wi.end(), nt);
    } else {
        const int64_t r = rays_override >= 0 ? rays_override : o->rays_per_emitter;
        const int64_t wi = (r + 3) / 4;
        if (o->rng == RF_RNG_REFERENCE && wi > ((int64_t)1 << 32))
            return set_err(ctx, RF_ERR_OVERFLOW, "reference RNG keys work items with 32 bits");
        std::fill(n_wi.begin(), n_wi.end(), wi);
    }
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Dispatch
// ---------------------------------------------------------------------------
static constexpr int kPendK = 4;            // deferred FP64 re-solves per lane (PEND_K in the kernels)
static constexpr int kMaxLdsBins = 15000;   // 32-bit bins per workgroup (< 64 KiB with the re-solve list)
static constexpr int kBnr = 4;              // binned BV form: rays per thread (BNR in the kernels)
static constexpr int kBsub = 4;             // binned BV form: sub-passes per bucketing (BSUB)
static constexpr size_t kMaxLds = 64 * 1024;

// Dynamic LDS of the binned BV kernel without / with tally bins.
static size_t bvbin_lds(int block, int bv_slots, int bins)
{
    return (size_t)kPendK * block * 4 + (size_t)block * kBnr * kBsub * 4 + (size_t)4 * bv_slots * 4 +
           (size_t)bins * 4;
}

// searched: some scene of the solve has a band stack to search (the untuned
// default shape depends on it).
static void pick_shape(const rf_ctx *ctx, const rf_solve_opts *o, int &nr, int &block, int &mfma,
                       bool searched = false)
{
    nr = 1;
    block = 256;
    mfma = 0;
    if (o->math == RF_MATH_FAST && !o->use_bvs) {
        const int tr = o->transform ? o->transform : (ctx->tuned_transform ? ctx->tuned_transform : RF_TRANSFORM_VALU);
        mfma = tr == RF_TRANSFORM_MFMA;
        // untuned default: 2 rays per thread where band stacks are searched
        // (the search is compiled for 1 and 2), else 4
        nr = o->rays_per_thread ? o->rays_per_thread
                                : (ctx->tuned_rpt ? ctx->tuned_rpt : (o->per_test == 0 && searched ? 2 : 4));
        block = o->block_size ? o->block_size : (ctx->tuned_block ? ctx->tuned_block : 256);
        if (mfma) nr = 4;   // 64 rays per wave and pass: one work item per lane-quarter
    } else if (o->block_size) {
        block = o->block_size;
    }
}

// Band stacks run in the fast form with the VALU transform (plain scan, and
// the children of the binned BV kernel) and in the compat plain scan, unless
// the caller asks for every test in full.
static bool uses_stacks(const rf_ctx *ctx, const rf_solve_opts *o)
{
    int nr, block, mfma;
    pick_shape(ctx, o, nr, block, mfma);
    if (o->per_test == 1) return false;
    return (o->math == RF_MATH_FAST && !mfma) || o->math == RF_MATH_COMPAT || o->math == RF_MATH_FP64;
}

// Records of a scene's fast-form cylinder segment (the band records the
// searched stacks read).
static int32_t cyl_records(const rf_scene *s) { return s->desc.seg_end[SEG_CYL] - s->desc.seg_begin[SEG_CYL]; }

// Launch switches of the fast band search (1 and 2 rays per thread, VALU
// transform, sorted stacks searched): the device counters (count_tests), and
// the LDS copy of each unit's band records when every scene's cylinder
// segment fits the tile (band_tile: -1 off, 0 the tuned choice -- off when
// untuned: on the 256-geometry sweep the copy measured 1.2% slower than the
// cached global reads -- else the most records staged) and the group's LDS
// stays within kMaxLds.  Other launches keep the counters and read the
// records from global memory.
static void band_switches(const rf_ctx *ctx, const rf_solve_opts *o, LaunchCfg &c, const KParams &kp,
                          int32_t max_cyl, bool searched)
{
    c.count = 1;
    c.bt = 0;
    const bool search = c.math == KM_FAST && !c.use_bv && !c.trace && !c.direct && !c.mfma && c.nr <= 2 &&
                        kp.stacks == 2 && searched;
    if (!search) return;
    c.count = o->count_tests;
    const bool on = o->band_tile > 0 || (o->band_tile == 0 && ctx->tuned_tile == 2);
    const int32_t cap = o->band_tile > 0 ? o->band_tile : (1 << 13);
    const size_t bytes = ((c.lds_bytes + 7) & ~(size_t)7) + (size_t)max_cyl * 8;
    if (on && max_cyl <= cap && bytes <= kMaxLds) {
        c.bt = 1;
        c.lds_bytes = bytes;
    }
}

// Whether a solve's units of scene s walk its hierarchy: the fast form where
// the scene has one, the FP64 form where it holds every primitive (no band
// stacks); the hierarchy's depth if so, else 0.
static int walked_depth(const rf_scene *s, const rf_solve_opts *o)
{
    if (o->math == RF_MATH_FP64) return s->desc.bvh_on == 2 ? s->bvh_depth : 0;
    return o->math == RF_MATH_FAST && s->bvh_on ? s->bvh_depth : 0;
}

// Rays per thread of the hierarchy form (compiled for 1, 2, 4 and 8) unless
// the caller asks: 4 where no scene of the launch has band stacks (with
// half-precision nodes and four sub-passes per tile, 4096-ray tiles measured
// 1.03x to 1.3x 2 rays per thread on 300- to 2000-primitive soups); 2 where
// one has, since band stacks are searched only at 1 and 2.
static constexpr int kBvhRpt = 4;
static constexpr int kBvhRptStacks = 2;

// The hierarchy form of the fast plain scan (VALU transform, band stacks on,
// the caller's hierarchy switch on) when some scene of the launch has a
// hierarchy (depth: the deepest one's; the units of the others scan):
// kBvhRpt (kBvhRptStacks) rays per thread unless asked, no LDS band copy, the device counters
// as count_tests asks, and each lane's traversal stack (depth + 1 entries)
// after the re-solve list (the LDS tally is dropped if they would not fit).
static void bvh_switch(const rf_solve_opts *o, LaunchCfg &c, KParams &kp, int depth, bool stacks)
{
    c.bvh = 0;
    if (depth <= 0 || o->hierarchy != 0 || (c.math != KM_FAST && c.math != KM_FP64) || c.use_bv || c.direct ||
        c.mfma || !kp.stacks)
        return;
    // (the FP64 walk's tiles are of 2 rays per thread whatever the option)
    int nr = c.math == KM_FP64 ? 2 : (o->rays_per_thread ? o->rays_per_thread : stacks ? kBvhRptStacks : kBvhRpt);
    const size_t fixed = (size_t)(kPendK + depth + 1) * c.block * sizeof(uint32_t);
    // the kernel's own LDS: the tile sort's cell counts and two 16-bit words
    // per ray of a tile (four sub-passes of up to 256 x nr rays); fewer rays
    // per thread where a deep hierarchy leaves no room, the scan if none fits
    const auto sort = [](int r) { return 512 * sizeof(uint32_t) + (size_t)4 * 256 * r * 2 * sizeof(uint16_t); };
    while (c.math == KM_FAST && nr > 1 && fixed + sort(nr) > kMaxLds) nr /= 2;
    if (fixed + sort(nr) > kMaxLds) return;
    c.bvh = 1;
    c.bt = 0;
    c.count = c.math == KM_FP64 ? 1 : o->count_tests;   // the FP64 walk always counts
    c.nr = nr;
    if (!c.trace && kp.bins && fixed + sort(nr) + (size_t)kp.bins * sizeof(uint32_t) > kMaxLds) kp.bins = 0;
    c.lds_bytes = fixed + (c.trace ? 0 : (size_t)kp.bins * sizeof(uint32_t));
}

static int ensure_queue(rf_ctx *ctx)
{
    if (ctx->d_queue) return RF_OK;
    HIPCHK(ctx, hipMalloc((void **)&ctx->d_queue, kFan * kRing * sizeof(unsigned)));
    return RF_OK;
}

static hipStream_t lane_stream(const rf_ctx *ctx, int lane) { return lane == 0 ? ctx->stream : ctx->fan[lane]; }

// One queue counter of a lane's ring (not yet zeroed).
static int take_counter(rf_ctx *ctx, unsigned **slot, int lane = 0)
{
    int rc = ensure_queue(ctx);
    if (rc) return rc;
    *slot = ctx->d_queue + (size_t)lane * kRing + (ctx->queue_next[lane]++ % kRing);
    return RF_OK;
}

// One queue counter, zeroed on the lane's stream right before the launch that uses it.
static int next_counter(rf_ctx *ctx, unsigned **slot, int lane)
{
    int rc = take_counter(ctx, slot, lane);
    if (rc) return rc;
    HIPCHK(ctx, launch_fill_u32(*slot, 0u, 1, lane_stream(ctx, lane)));
    return RF_OK;
}

// Side lanes start after everything queued on `stream` so far.
static int fan_fork(rf_ctx *ctx)
{
    if (ctx->fan_open) return RF_OK;
    HIPCHK(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
    for (int i = 1; i < kFan; ++i) HIPCHK(ctx, hipStreamWaitEvent(ctx->fan[i], ctx->fork_ev, 0));
    ctx->fan_open = true;
    ctx->fan_next = 0;
    return RF_OK;
}

// `stream` continues after everything queued on the side lanes.
static int fan_join(rf_ctx *ctx)
{
    if (!ctx->fan_open) return RF_OK;
    for (int i = 1; i < kFan; ++i) {
        HIPCHK(ctx, hipEventRecord(ctx->fan_ev[i], ctx->fan[i]));
        HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->fan_ev[i], 0));
    }
    ctx->fan_open = false;
    return RF_OK;
}

static int fan_lane(rf_ctx *ctx)
{
    const int l = ctx->fan_next;
    ctx->fan_next = (l + 1) % kFan;
    return l;
}

// Work units per launched group in a persistent solve: the launch's tail is
// about half a unit per group, so finer units balance better.  32 measured
// 10% faster than 8 on a 16-geometry sweep sample at 1e6 rays per emitter
// (no change on the 256-geometry sweep, whose emitters outnumber the units).
static constexpr double kUnitsPerGroup = 32.0;

static int max_groups(const rf_ctx *ctx, int block)
{
    // at least 4 resident groups per CU keeps every SIMD busy with 256-thread groups
    const int per_cu = std::max(2, 2048 / block);
    return ctx->n_cu * per_cu;
}

// Implicit-unit launch over work items [wi0, wi0+n_wi) of one emitter.
// lane: fan lane whose stream takes the launch (0 = the context stream).
static int launch_range(rf_ctx *ctx, rf_scene *s, int emitter, int64_t wi0, int64_t n_wi,
                        const rf_solve_opts *o, uint32_t geom_id, int32_t *trace_out, int64_t trace_wi0,
                        int64_t *launches, int lane = 0);

static int launch_cfg_solve(rf_ctx *ctx, const LaunchCfg &c, const KParams &kp, hipStream_t st)
{
    if (!c.trace) ctx->last_cfg = c;
    HIPCHK(ctx, launch_solve(c, kp, st));
    return RF_OK;
}

static int launch_range(rf_ctx *ctx, rf_scene *s, int emitter, int64_t wi0, int64_t n_wi,
                        const rf_solve_opts *o, uint32_t geom_id, int32_t *trace_out, int64_t trace_wi0,
                        int64_t *launches, int lane)
{
    const hipStream_t st = lane_stream(ctx, lane);
    if (n_wi <= 0) return RF_OK;
    int nr, block, mfma;
    pick_shape(ctx, o, nr, block, mfma, s->has_search);
    KParams kp{};
    kp.geoms = s->d_desc;
    kp.units = nullptr;
    kp.stats = ctx->d_stats;
    kp.trace_out = trace_out;
    kp.trace_wi0 = trace_wi0;
    kp.rng = o->rng;
    kp.seed = o->seed;
    kp.refine = o->refine;
    kp.tau = o->refine_margin > 0.0f ? o->refine_margin : rfd::TAU_DEFAULT;
    kp.bins = s->n <= kMaxLdsBins ? s->n : 0;
    kp.stacks = !uses_stacks(ctx, o) ? 0 : (o->per_test == 2 ? 1 : 2);
    kp.imp_emitter = emitter;
    kp.imp_geom_id = geom_id;
    kp.imp_wi0 = wi0;
    kp.imp_n_wi = n_wi;
    LaunchCfg c{};
    c.math = o->math == RF_MATH_COMPAT ? KM_COMPAT : (o->math == RF_MATH_FP64 ? KM_FP64 : KM_FAST);
    c.use_bv = o->use_bvs && s->n_bvs > 0;
    c.trace = trace_out != nullptr;
    c.nr = nr;
    c.refine = o->refine;
    c.mfma = mfma;
    if (o->form == RF_FORM_REFERENCE && !trace_out) {
        // main.cpp:172-176: global = nthreads, local = 128
        c.direct = 1;
        c.block = 128;
        const int64_t grid = (n_wi + 127) / 128;
        if (grid > INT32_MAX) return set_err(ctx, RF_ERR_CAPACITY, "grid too large");
        c.grid = (int)grid;
        c.lds_bytes = (size_t)kPendK * c.block * sizeof(uint32_t);
        const int rcd = launch_cfg_solve(ctx, c, kp, st);
        if (rcd) return rcd;
        ++*launches;
        return RF_OK;
    }
    // queue form: units of at least one pass of the workgroup, sized to give
    // every resident group about one unit
    const int64_t pass_wi = std::max<int64_t>(1, (int64_t)block * (nr >= 4 ? nr / 4 : 1));
    const int groups = max_groups(ctx, block);
    int64_t unit = (n_wi + groups - 1) / groups;
    unit = std::min<int64_t>(unit, (int64_t)1 << 28);   // 32-bit LDS bins: < 2^30 rays per unit
    unit = ((unit + pass_wi - 1) / pass_wi) * pass_wi;
    unit = std::max(unit, pass_wi);
    const int64_t n_units = (n_wi + unit - 1) / unit;
    if (n_units > INT32_MAX) return set_err(ctx, RF_ERR_CAPACITY, "ray budget too large for one launch");
    kp.imp_unit = unit;
    kp.n_units = (int32_t)n_units;
    int rc = next_counter(ctx, &kp.queue_head, lane);
    if (rc) return rc;
    c.direct = 0;
    c.block = block;
    c.grid = (int)std::min<int64_t>(n_units, groups);
    c.lds_bytes = ((c.trace || kp.bins == 0) ? 0 : (size_t)kp.bins * sizeof(uint32_t)) +
                  (size_t)kPendK * block * sizeof(uint32_t);
    band_switches(ctx, o, c, kp, cyl_records(s), s->has_search);
    bvh_switch(o, c, kp, walked_depth(s, o), s->stack_prims > 0);
    if (c.math == KM_FAST && c.use_bv && s->desc.bv_binned) {
        c.bvbin = 1;
        kp.bv_slots = s->n_bvs;
        if (bvbin_lds(block, kp.bv_slots, kp.bins) > kMaxLds) kp.bins = 0;
        c.lds_bytes = bvbin_lds(block, kp.bv_slots, c.trace ? 0 : kp.bins);
    }
    rc = launch_cfg_solve(ctx, c, kp, st);
    if (rc) return rc;
    ++*launches;
    return RF_OK;
}

extern "C" int32_t rf_launch_emitter(rf_ctx *ctx, rf_scene *s, int32_t emitter, int64_t wi0, int64_t n_wi,
                                     const rf_solve_opts *o)
{
    if (!ctx || !s) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL context or scene");
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    if (emitter < 0 || emitter >= s->n) return set_err(ctx, RF_ERR_INVALID_ARG, "emitter out of range");
    if (wi0 < 0 || n_wi < 0) return set_err(ctx, RF_ERR_INVALID_ARG, "negative work-item range");
    if (o->rng == RF_RNG_REFERENCE && wi0 + n_wi > ((int64_t)1 << 32))
        return set_err(ctx, RF_ERR_OVERFLOW, "reference RNG keys work items with 32 bits");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int64_t launches = 0;
    // the reference shape keeps one in-order queue; queue-kernel launches fan out
    if (o->form == RF_FORM_REFERENCE) {
        rc = fan_join(ctx);
        if (rc) return rc;
        return launch_range(ctx, s, emitter, wi0, n_wi, o, o->geometry, nullptr, 0, &launches, 0);
    }
    rc = fan_fork(ctx);
    if (rc) return rc;
    return launch_range(ctx, s, emitter, wi0, n_wi, o, o->geometry, nullptr, 0, &launches, fan_lane(ctx));
}

extern "C" int32_t rf_sync(rf_ctx *ctx)
{
    if (!ctx) return RF_ERR_INVALID_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = fan_join(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RF_OK;
}

static int read_hits(rf_ctx *ctx, rf_scene *s, int64_t *hits)
{
    int rc = fan_join(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(hits, s->d_hits, (size_t)s->n * s->n * sizeof(int64_t), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RF_OK;
}

extern "C" int32_t rf_vf_read(rf_ctx *ctx, rf_scene *s, int64_t rays_per_emitter, double *vf)
{
    if (!ctx || !s || !vf) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    if (rays_per_emitter < 0) return set_err(ctx, RF_ERR_INVALID_ARG, "negative ray count");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = fan_join(ctx);
    if (rc) return rc;
    const size_t n2 = (size_t)s->n * s->n;
    double *d_vf = nullptr;
    HIPCHK(ctx, hipMalloc((void **)&d_vf, n2 * sizeof(double)));
    hipError_t e = launch_normalise(s->d_hits, d_vf, n2, (double)rays_per_emitter, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vf, d_vf, n2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_vf);
    if (e != hipSuccess) return set_err(ctx, RF_ERR_HIP, std::string("view-factor normalisation: ") + hipGetErrorString(e));
    return RF_OK;
}

extern "C" int32_t rf_hits_read(rf_ctx *ctx, rf_scene *s, int64_t *hits)
{
    if (!ctx || !s || !hits) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return read_hits(ctx, s, hits);
}

extern "C" int32_t rf_hits_reset(rf_ctx *ctx, rf_scene *s)
{
    if (!ctx || !s) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = fan_join(ctx);
    if (rc) return rc;
    HIPCHK(ctx, launch_fill_u64(s->d_hits, (size_t)s->n * s->n, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RF_OK;
}

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int read_stats(rf_ctx *ctx, unsigned long long st[kNStats])
{
    HIPCHK(ctx, hipMemcpy(st, ctx->d_stats, kNStats * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return RF_OK;
}

// Accounting of one geometry's solve.
static void account(const rf_ctx *ctx, const rf_scene *s, const rf_solve_opts *o, const std::vector<int64_t> &n_wi,
                    rf_timing &t)
{
    int64_t rays = 0;
    for (int64_t w : n_wi) rays += 4 * w;
    t.rays += rays;
    t.tests_brute_force += rays * s->n;
    if (!(o->use_bvs && s->n_bvs > 0)) {
        // band stacks add their evaluations and flops from the device
        // counters (stats[4], stats[3]); every other primitive is one test
        const bool st = uses_stacks(ctx, o);
        int nr, block, mfma;
        pick_shape(ctx, o, nr, block, mfma);
        const bool fast_valu = o->math == RF_MATH_FAST && !mfma;   // scan_fast: world z-plane rings
        t.tests += rays * (!st ? s->n : (o->math == RF_MATH_FAST ? s->n_nonstack : s->n_cnonstack));
        const int64_t f = !st ? (fast_valu ? s->flops_per_ray_fast : s->flops_per_ray)
                          : (o->math == RF_MATH_FAST ? s->flops_per_ray_nonstack : s->flops_per_ray_cnonstack);
        t.flops += rays * f;
    }
}

// Persistent batched solve over several scenes (and the single-scene
// persistent form).
// Units, descriptors and launch configuration of a persistent solve; the
// descriptors and units are uploaded to d_geoms / d_units (stream order), the
// queue counter is chosen but not zeroed (enqueue_persistent does that).
static int prepare_persistent(rf_ctx *ctx, rf_scene *const *scenes, int n_geoms, const rf_solve_opts *o,
                              const std::vector<std::vector<int64_t>> &budgets, const std::vector<uint32_t> &ids,
                              LaunchCfg &c, KParams &kp, bool &empty)
{
    bool searched = false;
    for (int g = 0; g < n_geoms; ++g) searched = searched || scenes[g]->has_search;
    int nr, block, mfma;
    pick_shape(ctx, o, nr, block, mfma, searched);
    const bool use_bv = o->use_bvs != 0;
    // geometry descriptors
    std::vector<GeomDesc> descs(n_geoms);
    int max_n = 0;
    for (int g = 0; g < n_geoms; ++g) {
        descs[g] = scenes[g]->desc;
        descs[g].geometry_id = ids[g];
        max_n = std::max(max_n, scenes[g]->n);
    }
    const int groups = max_groups(ctx, block);
    const int64_t pass_wi = std::max<int64_t>(1, (int64_t)block * (nr >= 4 ? nr / 4 : 1));
    // Units by cost (work items x primitives): about kUnitsPerGroup per
    // launched group for balance.  An emitter is split only when it costs more
    // than that, so most emitter rows are flushed from LDS a few times at most
    // (one 64-bit atomic per bin).
    double total_cost = 0.0;
    for (int g = 0; g < n_geoms; ++g) total_cost += (double)scenes[g]->n * (double)std::accumulate(
                                                        budgets[g].begin(), budgets[g].end(), (int64_t)0);
    const double unit_cost = std::max(1.0, total_cost / ((double)groups * kUnitsPerGroup));
    const int64_t max_unit_wi = (int64_t)1 << 28;   // 32-bit LDS bins: < 2^30 rays per unit
    std::vector<WorkUnit> units;
    for (int g = 0; g < n_geoms; ++g)
        for (int e = 0; e < scenes[g]->n; ++e) {
            const int64_t nw = budgets[g][e];
            if (nw <= 0) continue;
            const int64_t pieces = std::max<int64_t>(
                (int64_t)std::ceil((double)nw * scenes[g]->n / unit_cost), (nw + max_unit_wi - 1) / max_unit_wi);
            int64_t size = (nw + pieces - 1) / pieces;
            size = ((size + pass_wi - 1) / pass_wi) * pass_wi;
            for (int64_t w0 = 0; w0 < nw; w0 += size)
                units.push_back(WorkUnit{g, e, w0, std::min(size, nw - w0), 0});
        }
    empty = units.empty();
    if (empty) return RF_OK;
    if (units.size() > (size_t)INT32_MAX) return set_err(ctx, RF_ERR_CAPACITY, "too many work units");
    // largest units first: cost ~ work items x primitives
    std::stable_sort(units.begin(), units.end(), [&](const WorkUnit &a, const WorkUnit &b) {
        return a.n_wi * scenes[a.geom]->n > b.n_wi * scenes[b.geom]->n;
    });
    if (ctx->geoms_cap < (size_t)n_geoms) {
        (void)hipFree(ctx->d_geoms);
        HIPCHK(ctx, hipMalloc((void **)&ctx->d_geoms, n_geoms * sizeof(GeomDesc)));
        ctx->geoms_cap = n_geoms;
    }
    if (ctx->units_cap < units.size()) {
        (void)hipFree(ctx->d_units);
        HIPCHK(ctx, hipMalloc((void **)&ctx->d_units, units.size() * sizeof(WorkUnit)));
        ctx->units_cap = units.size();
    }
    // the host vectors die with this call: copy synchronously (before any launch)
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_geoms, descs.data(), n_geoms * sizeof(GeomDesc), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_units, units.data(), units.size() * sizeof(WorkUnit), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    kp = KParams{};
    kp.geoms = ctx->d_geoms;
    kp.units = ctx->d_units;
    kp.stats = ctx->d_stats;
    kp.n_units = (int32_t)units.size();
    kp.rng = o->rng;
    kp.seed = o->seed;
    kp.refine = o->refine;
    kp.tau = o->refine_margin > 0.0f ? o->refine_margin : rfd::TAU_DEFAULT;
    kp.bins = max_n <= kMaxLdsBins ? max_n : 0;
    kp.stacks = !uses_stacks(ctx, o) ? 0 : (o->per_test == 2 ? 1 : 2);
    int rc = take_counter(ctx, &kp.queue_head);
    if (rc) return rc;
    c = LaunchCfg{};
    c.math = o->math == RF_MATH_COMPAT ? KM_COMPAT : (o->math == RF_MATH_FP64 ? KM_FP64 : KM_FAST);
    c.use_bv = use_bv;
    c.nr = nr;
    c.refine = o->refine;
    c.mfma = mfma;
    c.block = block;
    c.grid = (int)std::min<int64_t>((int64_t)units.size(), groups);
    c.lds_bytes = (kp.bins ? (size_t)kp.bins * sizeof(uint32_t) : 0) + (size_t)kPendK * block * sizeof(uint32_t);
    int32_t max_cyl = 0;
    for (int g = 0; g < n_geoms; ++g) max_cyl = std::max(max_cyl, cyl_records(scenes[g]));
    band_switches(ctx, o, c, kp, max_cyl, searched);
    int bvh_depth = 0;
    bool any_stacks = false;
    for (int g = 0; g < n_geoms; ++g) {
        bvh_depth = std::max(bvh_depth, walked_depth(scenes[g], o));
        any_stacks = any_stacks || scenes[g]->stack_prims > 0;
    }
    bvh_switch(o, c, kp, bvh_depth, any_stacks);
    if (c.math == KM_FAST && use_bv) {
        bool binned = true;
        int slots = 0;
        for (int g = 0; g < n_geoms; ++g) {
            binned = binned && scenes[g]->desc.bv_binned;
            slots = std::max(slots, scenes[g]->n_bvs);
        }
        if (binned) {
            c.bvbin = 1;
            kp.bv_slots = slots;
            if (bvbin_lds(block, slots, kp.bins) > kMaxLds) kp.bins = 0;
            c.lds_bytes = bvbin_lds(block, slots, kp.bins);
        }
    }
    return RF_OK;
}

// Zero fills of the tallies, statistics and queue counter, then the launch.
// All kernels (no hipMemsetAsync), so the sequence can be captured in a graph.
static hipError_t enqueue_persistent(rf_ctx *ctx, rf_scene *const *scenes, int n_geoms, const LaunchCfg &c,
                                     const KParams &kp)
{
    hipError_t e = hipSuccess;
    if (kp.geoms) {
        // one launch: the uploaded descriptors kp.geoms hold each scene's tally (d_hits)
        size_t cells = 0;
        for (int g = 0; g < n_geoms; ++g) cells += (size_t)scenes[g]->n * scenes[g]->n;
        e = launch_fill_solve(kp.geoms, n_geoms, cells, ctx->d_stats, kp.queue_head, ctx->stream);
    } else {
        // an empty solve uploads no descriptors (prepare_persistent): fill each scene's tally
        for (int g = 0; g < n_geoms && e == hipSuccess; ++g)
            e = launch_fill_u64(scenes[g]->d_hits, (size_t)scenes[g]->n * scenes[g]->n, ctx->stream);
        if (e == hipSuccess) e = launch_fill_u64(ctx->d_stats, kNStats, ctx->stream);
        if (e == hipSuccess && kp.queue_head) e = launch_fill_u32(kp.queue_head, 0u, 1, ctx->stream);
    }
    if (e == hipSuccess && kp.n_units > 0) e = launch_solve(c, kp, ctx->stream);
    return e;
}

// Identity of a persistent solve for the graph cache: scene objects, options,
// launch shape choices, geometry ids and ray budgets.
static std::vector<uint64_t> persistent_key(const rf_ctx *ctx, rf_scene *const *scenes, int n_geoms,
                                            const rf_solve_opts *o, const std::vector<uint32_t> &ids,
                                            const std::vector<std::vector<int64_t>> &budgets)
{
    std::vector<uint64_t> k;
    k.push_back((uint64_t)n_geoms);
    unsigned char ob[sizeof(rf_solve_opts)];
    std::memcpy(ob, o, sizeof ob);
    for (size_t i = 0; i < sizeof ob; i += 8) {
        uint64_t w = 0;
        std::memcpy(&w, ob + i, std::min<size_t>(8, sizeof ob - i));
        k.push_back(w);
    }
    k.push_back(((uint64_t)ctx->tuned_rpt << 40) ^ ((uint64_t)ctx->tuned_block << 20) ^
                ((uint64_t)ctx->tuned_tile << 8) ^ (uint64_t)ctx->tuned_transform);
    for (int g = 0; g < n_geoms; ++g) {
        k.push_back(scenes[g]->serial);
        k.push_back(ids[g]);
        uint64_t h = 1469598103934665603ull;   // FNV-1a over the per-emitter budgets
        for (int64_t w : budgets[g]) { h ^= (uint64_t)w; h *= 1099511628211ull; }
        k.push_back(h);
        k.push_back(budgets[g].size());
    }
    return k;
}

static void drop_graph(rf_ctx *ctx)
{
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_key.clear();
}

// Persistent solve between the timing events: replay the cached graph when
// the same solve ran before; otherwise prepare, and run eagerly (first time)
// or capture and launch a graph (second time).
static int solve_persistent(rf_ctx *ctx, rf_scene *const *scenes, int n_geoms, const rf_solve_opts *o,
                            const std::vector<std::vector<int64_t>> &budgets, const std::vector<uint32_t> &ids,
                            int64_t *launches, int64_t *graphed)
{
    const std::vector<uint64_t> key = persistent_key(ctx, scenes, n_geoms, o, ids, budgets);
    if (ctx->graph_exec && key == ctx->graph_key) {
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
        ++ctx->graph_replays;
        ++*launches;
        ++*graphed;
        return RF_OK;
    }
    // d_geoms / d_units are about to change: the cached graph no longer applies
    drop_graph(ctx);
    LaunchCfg c{};
    KParams kp{};
    bool empty = false;
    int rc = prepare_persistent(ctx, scenes, n_geoms, o, budgets, ids, c, kp, empty);
    if (rc) return rc;
    ctx->last_cfg = c;
    if (!empty && key == ctx->last_key) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue_persistent(ctx, scenes, n_geoms, c, kp);
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return set_err(ctx, RF_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        const hipError_t e3 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_err(ctx, RF_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e3));
        }
        ctx->graph = g;
        ctx->graph_exec = ge;
        ctx->graph_key = key;
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(ctx, hipGraphLaunch(ge, ctx->stream));
        ++*graphed;
    } else {
        ctx->last_key = key;
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(ctx, enqueue_persistent(ctx, scenes, n_geoms, c, kp));
    }
    if (!empty) ++*launches;
    return RF_OK;
}

// geom_ids: seeded-RNG geometry word of each scene (NULL: opts->geometry + index);
// geom_rays: explicit rays per emitter of each scene (NULL: opts->rays_per_emitter)
// geom_emitters: [lo, hi) of the emitters of each scene that get rays (NULL:
// all; the others get none and their rows stay zero)
static int solve_impl(rf_ctx *ctx, rf_scene *const *scenes, int n_geoms, const rf_solve_opts *o,
                      int64_t *const *hits, int64_t *const *rays, rf_timing *timing,
                      const uint32_t *geom_ids = nullptr, const int64_t *geom_rays = nullptr,
                      const int32_t (*geom_emitters)[2] = nullptr)
{
    const double w0 = now_ms();
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    if (ctx->fail_in > 0 && --ctx->fail_in == 0)
        return set_err(ctx, RF_ERR_HIP, "injected failure (rf_debug_fail_solve)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<std::vector<int64_t>> budgets(n_geoms);
    std::vector<uint32_t> ids(n_geoms);
    for (int g = 0; g < n_geoms; ++g) ids[g] = geom_ids ? geom_ids[g] : o->geometry + (uint32_t)g;
    for (int g = 0; g < n_geoms; ++g) {
        if (!scenes[g] || scenes[g]->ctx != ctx) return set_err(ctx, RF_ERR_INVALID_ARG, "scene of another context");
        rc = emitter_budget(ctx, scenes[g], o, budgets[g], geom_rays ? geom_rays[g] : -1);
        if (rc) return rc;
        if (geom_emitters)
            for (int e = 0; e < scenes[g]->n; ++e)
                if (e < geom_emitters[g][0] || e >= geom_emitters[g][1]) budgets[g][e] = 0;
    }
    rc = fan_join(ctx);
    if (rc) return rc;
    rc = ensure_queue(ctx);
    if (rc) return rc;
    int64_t launches = 0, graphed = 0;
    if (o->form == RF_FORM_PERSISTENT) {
        // zero fills are part of the (graph-captured) persistent sequence
        rc = solve_persistent(ctx, scenes, n_geoms, o, budgets, ids, &launches, &graphed);
    } else {
        // hit matrices and statistics start at zero
        for (int g = 0; g < n_geoms; ++g)
            HIPCHK(ctx, launch_fill_u64(scenes[g]->d_hits, (size_t)scenes[g]->n * scenes[g]->n, ctx->stream));
        HIPCHK(ctx, launch_fill_u64(ctx->d_stats, kNStats, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        // one launch per emitter; queue-kernel launches of different emitters
        // are independent and fan out over the lanes, the reference shape
        // keeps one in-order queue
        const bool fan = o->form != RF_FORM_REFERENCE;
        if (fan) rc = fan_fork(ctx);
        for (int g = 0; g < n_geoms && !rc; ++g)
            for (int e = 0; e < scenes[g]->n && !rc; ++e)
                rc = launch_range(ctx, scenes[g], e, 0, budgets[g][e], o, ids[g], nullptr, 0, &launches,
                                  fan ? fan_lane(ctx) : 0);
        const int rj = fan_join(ctx);
        if (!rc) rc = rj;
    }
    if (rc) { (void)hipStreamSynchronize(ctx->stream); return rc; }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long st[kNStats];
    rc = read_stats(ctx, st);
    if (rc) return rc;
    rf_timing t{};
    t.kernel_ms = ms;
    t.launches = launches;
    t.graph_launches = graphed;
    t.refined = (int64_t)st[1];
    t.rescans = (int64_t)st[2];
    for (int g = 0; g < n_geoms; ++g) {
        account(ctx, scenes[g], o, budgets[g], t);
        if (o->use_bvs && scenes[g]->n_bvs > 0) {
            // BV kernels count executed tests on the device (all geometries together)
        }
        if (hits && hits[g]) { rc = read_hits(ctx, scenes[g], hits[g]); if (rc) return rc; }
        if (rays && rays[g])
            for (int e = 0; e < scenes[g]->n; ++e) rays[g][e] = 4 * budgets[g][e];
    }
    if (!o->use_bvs && uses_stacks(ctx, o)) {
        t.flops += (int64_t)st[3];
        t.tests += (int64_t)st[4];
        // the fast band search ran without its counters: not measured
        int nr, block, mfma;
        bool searched = false;
        for (int g = 0; g < n_geoms; ++g) searched = searched || scenes[g]->has_search;
        pick_shape(ctx, o, nr, block, mfma, searched);
        if (o->math == RF_MATH_FAST && !o->count_tests && o->per_test == 0 && searched && nr <= 2 && !mfma)
            t.tests = t.flops = -1;
    }
    if (o->use_bvs) {
        t.tests += (int64_t)st[0];
        // the binned kernel counts its children's flops on the device (stats[3],
        // stacked or per test); the surfaces' own tests are rays x sum of F
        bool binned = o->math == RF_MATH_FAST && o->form != RF_FORM_REFERENCE;
        for (int g = 0; g < n_geoms; ++g) binned = binned && scenes[g]->n_bvs > 0 && scenes[g]->desc.bv_binned;
        if (binned) {
            t.flops += (int64_t)st[3];
            for (int g = 0; g < n_geoms; ++g)
                for (int64_t w : budgets[g]) t.flops += 4 * w * scenes[g]->flops_bvs;
        } else {
            double favg = 0.0;
            int nbv_g = 0;
            for (int g = 0; g < n_geoms; ++g)
                if (scenes[g]->n_bvs > 0) { favg += scenes[g]->flops_per_test_avg; ++nbv_g; }
            favg = nbv_g ? favg / nbv_g : 0.0;
            t.flops += (int64_t)((double)st[0] * favg);
        }
    }
    const LaunchCfg &lc = ctx->last_cfg;
    if (lc.bvh) {
        // the hierarchy's box and primitive tests and the band stacks'
        // evaluations, counted on the device (or not at all); the per-test
        // primitives of scenes without a hierarchy, one test each
        t.tests = t.flops = -1;
        if (lc.count) {
            t.tests = (int64_t)st[4];
            t.flops = (int64_t)st[3];
            const bool fast = o->math == RF_MATH_FAST;
            for (int g = 0; g < n_geoms; ++g) {
                if (walked_depth(scenes[g], o)) continue;
                int64_t r = 0;
                for (int64_t w : budgets[g]) r += 4 * w;
                t.tests += r * (fast ? scenes[g]->n_nonstack : scenes[g]->n_cnonstack);
                t.flops += r * (fast ? scenes[g]->flops_per_ray_nonstack : scenes[g]->flops_per_ray_cnonstack);
            }
        }
    }
    t.shape_rays_per_thread = lc.mfma ? 0 : (lc.direct || o->math != RF_MATH_FAST || o->use_bvs ? 4 : lc.nr);
    t.shape_block = lc.block;
    t.shape_flags = (lc.mfma ? RF_SHAPE_MFMA : 0) | (lc.bvbin ? RF_SHAPE_BV_BINNED : 0) |
                    (lc.direct ? RF_SHAPE_REFERENCE : 0) | (lc.bt ? RF_SHAPE_BAND_TILE : 0) |
                    (uses_stacks(ctx, o) ? RF_SHAPE_STACKS : 0) |
                    (uses_stacks(ctx, o) && o->per_test == 0 && lc.math == KM_FAST && lc.nr <= 2 && !lc.mfma &&
                             !lc.use_bv
                         ? RF_SHAPE_SEARCHED : 0) |
                    (lc.count || !(lc.math == KM_FAST && lc.nr <= 2) ? RF_SHAPE_COUNTED : 0) |
                    (lc.bvh ? RF_SHAPE_HIERARCHY : 0);
    t.wall_ms = now_ms() - w0;
    if (timing) *timing = t;
    return RF_OK;
}

extern "C" int32_t rf_solve(rf_ctx *ctx, rf_scene *s, const rf_solve_opts *o, int64_t *hits, int64_t *rays,
                            rf_timing *timing)
{
    if (!ctx || !s) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL context or scene");
    rf_scene *const sc[1] = {s};
    int64_t *const h[1] = {hits};
    int64_t *const r[1] = {rays};
    return solve_impl(ctx, sc, 1, o, h, r, timing);
}

// rf_solve's persistent form without any allocation: one launch whose work
// units are implicit over every emitter (all emitters share one budget),
// tallying into the caller's device matrix; the queue counter and statistics
// are the device's scratch, cleared by fill kernels on the context stream
// before the launch.  Asynchronous; no host or device allocation, no
// hipMemsetAsync, nothing synchronised.
extern "C" int32_t rf_solve_noalloc(rf_ctx *ctx, rf_scene *s, const rf_solve_opts *o, uint64_t *d_hits)
{
    if (!ctx || !s || !d_hits) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL context, scene or tally");
    if (s->ctx != ctx) return set_err(ctx, RF_ERR_INVALID_ARG, "scene of another context");
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    if (o->form != RF_FORM_PERSISTENT)
        return set_err(ctx, RF_ERR_INVALID_ARG, "the allocation-free solve runs the persistent form only");
    if (!ctx->scratch_ptr) return set_err(ctx, RF_ERR_HIP, "scratch address not set up");
    // the budget every emitter shares (emitter_budget without its vector)
    int64_t n_wi;
    if (o->budget == RF_BUDGET_REFERENCE) {
        rc = rf_nthreads(s->prims[0].rayDensity, o->n_samples, 4, 128, &n_wi);
        if (rc) return set_err(ctx, rc, "rayDensity * n_samples overflows int32 (main.cpp:172)");
        if (n_wi > INT32_MAX) return set_err(ctx, RF_ERR_OVERFLOW, "work items exceed the reference's int gid");
    } else {
        n_wi = (o->rays_per_emitter + 3) / 4;
        if (o->rng == RF_RNG_REFERENCE && n_wi > ((int64_t)1 << 32))
            return set_err(ctx, RF_ERR_OVERFLOW, "reference RNG keys work items with 32 bits");
    }
    int nr, block, mfma;
    pick_shape(ctx, o, nr, block, mfma, s->has_search);
    // units: about kUnitsPerGroup per launched group over all emitters, whole
    // passes of the group, never spanning two emitters
    const int groups = max_groups(ctx, block);
    const int64_t pass_wi = std::max<int64_t>(1, (int64_t)block * (nr >= 4 ? nr / 4 : 1));
    const int64_t upe0 = std::max<int64_t>(1, ((int64_t)kUnitsPerGroup * groups + s->n - 1) / s->n);
    int64_t unit = std::max<int64_t>(1, (n_wi + upe0 - 1) / upe0);
    unit = std::min<int64_t>(unit, (int64_t)1 << 28);   // 32-bit LDS bins: < 2^30 rays per unit
    unit = std::max(pass_wi, (unit + pass_wi - 1) / pass_wi * pass_wi);
    const int64_t upe = n_wi > 0 ? (n_wi + unit - 1) / unit : 0;
    const int64_t n_units = upe * s->n;
    if (n_units > INT32_MAX || upe > INT32_MAX)
        return set_err(ctx, RF_ERR_CAPACITY,
                       "allocation-free solve: more than 2^31-1 work units for the scratch's 32-bit queue counter");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    rc = fan_join(ctx);
    if (rc) return rc;
    unsigned *queue = (unsigned *)ctx->scratch_ptr;
    unsigned long long *stats = (unsigned long long *)((unsigned char *)ctx->scratch_ptr + kScratchStatsOff);
    KParams kp{};
    kp.geoms = s->d_desc;
    kp.units = nullptr;
    kp.queue_head = queue;
    kp.stats = stats;
    kp.n_units = (int32_t)n_units;
    kp.rng = o->rng;
    kp.seed = o->seed;
    kp.refine = o->refine;
    kp.tau = o->refine_margin > 0.0f ? o->refine_margin : rfd::TAU_DEFAULT;
    kp.bins = s->n <= kMaxLdsBins ? s->n : 0;
    kp.stacks = !uses_stacks(ctx, o) ? 0 : (o->per_test == 2 ? 1 : 2);
    kp.imp_emitter = -1;
    kp.imp_upe = (int32_t)upe;
    kp.imp_geom_id = o->geometry;
    kp.imp_wi0 = 0;
    kp.imp_n_wi = n_wi;
    kp.imp_unit = unit;
    kp.hits = (unsigned long long *)d_hits;
    LaunchCfg c{};
    c.math = o->math == RF_MATH_COMPAT ? KM_COMPAT : (o->math == RF_MATH_FP64 ? KM_FP64 : KM_FAST);
    c.use_bv = o->use_bvs && s->n_bvs > 0;
    c.nr = nr;
    c.refine = o->refine;
    c.mfma = mfma;
    c.block = block;
    c.grid = (int)std::min<int64_t>(std::max<int64_t>(n_units, 1), groups);
    c.lds_bytes = (kp.bins ? (size_t)kp.bins * sizeof(uint32_t) : 0) + (size_t)kPendK * block * sizeof(uint32_t);
    band_switches(ctx, o, c, kp, cyl_records(s), s->has_search);
    bvh_switch(o, c, kp, walked_depth(s, o), s->stack_prims > 0);
    if (c.math == KM_FAST && c.use_bv && s->desc.bv_binned) {
        c.bvbin = 1;
        kp.bv_slots = s->n_bvs;
        if (bvbin_lds(block, kp.bv_slots, kp.bins) > kMaxLds) kp.bins = 0;
        c.lds_bytes = bvbin_lds(block, kp.bv_slots, kp.bins);
    }
    // the tally, the queue counter and the statistics start at zero (fill
    // kernels: the sequence stays capturable in a graph)
    HIPCHK(ctx, launch_fill_u64((unsigned long long *)d_hits, (size_t)s->n * s->n, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(queue, 0u, 1, ctx->stream));
    HIPCHK(ctx, launch_fill_u64(stats, kNStats, ctx->stream));
    if (n_units > 0) HIPCHK(ctx, launch_solve(c, kp, ctx->stream));
    return RF_OK;
}

extern "C" int32_t rf_batch_solve(rf_ctx *ctx, rf_scene *const *scenes, int32_t n_geoms, const rf_solve_opts *o,
                                  int64_t *const *hits, int64_t *const *rays, rf_timing *timing)
{
    if (!ctx || !scenes || n_geoms <= 0) return set_err(ctx, RF_ERR_INVALID_ARG, "empty batch");
    return solve_impl(ctx, scenes, n_geoms, o, hits, rays, timing);
}

extern "C" int32_t rf_trace_rays(rf_ctx *ctx, rf_scene *s, int32_t emitter, int64_t wi0, int64_t n_wi,
                                 const rf_solve_opts *o, int32_t *idx_out)
{
    if (!ctx || !s || !idx_out) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    if (emitter < 0 || emitter >= s->n || wi0 < 0 || n_wi <= 0) return set_err(ctx, RF_ERR_INVALID_ARG, "bad range");
    if (n_wi > ((int64_t)1 << 26)) return set_err(ctx, RF_ERR_CAPACITY, "trace window above 2^26 work items");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    rc = fan_join(ctx);
    if (rc) return rc;
    int32_t *d_out = nullptr;
    HIPCHK(ctx, hipMalloc((void **)&d_out, (size_t)n_wi * 4 * sizeof(int32_t)));
    HIPCHK(ctx, launch_fill_u32((unsigned *)d_out, 0xffffffffu, (size_t)n_wi * 4, ctx->stream));
    int64_t launches = 0;
    rc = launch_range(ctx, s, emitter, wi0, n_wi, o, o->geometry, d_out, wi0, &launches);
    if (rc == RF_OK) {
        hipError_t e = hipMemcpyAsync(idx_out, d_out, (size_t)n_wi * 4 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = set_err(ctx, RF_ERR_HIP, std::string("trace readback: ") + hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    return rc;
}

extern "C" int32_t rf_run_cl_gpurng(rf_ctx *ctx, const rf_primitive *prims, int32_t n, const rf_bounding_volume *bvs,
                                    int32_t nbv, int32_t use_bvs, int32_t n_samples, int32_t *hits,
                                    double *samples_for_obj, float *vfm)
{
    if (!ctx || !hits || !samples_for_obj) return set_err(ctx, RF_ERR_INVALID_ARG, "NULL argument");
    rf_scene *s = nullptr;
    int rc = rf_scene_create(ctx, prims, n, bvs, nbv, &s);
    if (rc) return rc;
    rf_solve_opts o;
    rf_solve_opts_default(&o);
    o.math = RF_MATH_COMPAT;
    o.use_bvs = use_bvs;
    o.budget = RF_BUDGET_REFERENCE;
    o.n_samples = n_samples;
    std::vector<int64_t> h((size_t)n * n), r(n);
    rc = rf_solve(ctx, s, &o, h.data(), r.data(), nullptr);
    rf_scene_destroy(s);
    if (rc) return rc;
    for (size_t i = 0; i < h.size(); ++i) {
        if (h[i] > INT32_MAX) return set_err(ctx, RF_ERR_OVERFLOW, "a hit count exceeds int32 (main.cpp:103)");
        hits[i] = (int32_t)h[i];
    }
    for (int i = 0; i < n; ++i) samples_for_obj[i] = (double)r[i];   // nthreads*count, main.cpp:178
    if (vfm)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                float v = 0.0f;
                v += (float)hits[(size_t)i * n + j];                  // main.cpp:231
                vfm[(size_t)i * n + j] = (float)((double)v / samples_for_obj[i]);   // :245
            }
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: fastest (rays per thread, block size) of the fast plain kernel
// on a bounded sample of the given scenes' work.
// ---------------------------------------------------------------------------
extern "C" int32_t rf_autotune(rf_ctx *ctx, rf_scene *const *scenes, int32_t n_geoms, const rf_solve_opts *o,
                               int32_t *best_rpt, int32_t *best_block)
{
    if (!ctx || !scenes || n_geoms <= 0) return set_err(ctx, RF_ERR_INVALID_ARG, "empty batch");
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    rf_solve_opts t = *o;
    t.math = RF_MATH_FAST;
    t.use_bvs = 0;
    t.form = RF_FORM_PERSISTENT;
    t.budget = RF_BUDGET_EXPLICIT;
    // sample: about 3e11 ray-primitive tests per candidate (a few device fills)
    double tests_per_budget_ray = 0.0;   // sum over geometries of emitters x primitives
    for (int g = 0; g < n_geoms; ++g) tests_per_budget_ray += (double)scenes[g]->n * scenes[g]->n;
    int64_t rays = (int64_t)(3.0e11 / std::max(1.0, tests_per_budget_ray));
    rays = std::max<int64_t>(rays, 4096);
    if (o->budget == RF_BUDGET_EXPLICIT && o->rays_per_emitter > 0) rays = std::min(rays, o->rays_per_emitter);
    t.rays_per_emitter = rays;
    t.count_tests = 0;                  // timed as the solves it tunes
    // candidates: VALU transform with 2 rays per thread (band records in LDS
    // or not: the scene tile), 4 and 8 rays per thread, MFMA transform
    const int cands[][3] = {{RF_TRANSFORM_VALU, 2, 1}, {RF_TRANSFORM_VALU, 2, -1}, {RF_TRANSFORM_VALU, 4, -1},
                            {RF_TRANSFORM_VALU, 8, -1}, {RF_TRANSFORM_MFMA, 4, -1}};
    const int blocks[] = {128, 256};
    double best = 1e300;
    int br = 4, bb = 256, bt = RF_TRANSFORM_VALU, btile = -1;
    for (const auto &cd : cands)
        for (int bl : blocks) {
            t.transform = cd[0];
            t.rays_per_thread = cd[1];
            t.band_tile = cd[2] > 0 ? (1 << 13) : -1;
            t.block_size = bl;
            rf_timing tm{};
            rc = solve_impl(ctx, scenes, n_geoms, &t, nullptr, nullptr, &tm);   // warm
            if (rc) return rc;
            rc = solve_impl(ctx, scenes, n_geoms, &t, nullptr, nullptr, &tm);
            if (rc) return rc;
            if (tm.kernel_ms < best) { best = tm.kernel_ms; br = cd[1]; bb = bl; bt = cd[0]; btile = cd[2]; }
        }
    ctx->tuned_transform = bt;
    if (bt == RF_TRANSFORM_MFMA) br = 0;
    ctx->tuned_rpt = br ? br : 4;
    ctx->tuned_block = bb;
    ctx->tuned_tile = btile > 0 ? 2 : 1;
    if (best_rpt) *best_rpt = br;
    if (best_block) *best_block = bb;
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Streaming service: resident scenes and tallies, admission by HBM footprint
// and by estimated cost, queued requests of unequal size packed shortest
// first into persistent launches.
// ---------------------------------------------------------------------------

// Two requests may share one persistent launch only if every option that can
// change a tally is equal: the arithmetic form, random stream, budget rule,
// FP64 re-solve and its margin, the transform unit and the scan form.  The
// launch shape (rays per thread, group size) is compared too, so a request
// runs in exactly the shape it asked for.  Per-request: the geometry word and
// the explicit ray budget.
static bool same_tallies(const rf_solve_opts &a, const rf_solve_opts &b)
{
    const auto margin = [](const rf_solve_opts &o) { return o.refine_margin > 0.0f ? o.refine_margin : 0x1p-14f; };
    const bool fast = a.math == RF_MATH_FAST;
    return a.math == b.math && a.rng == b.rng && a.use_bvs == b.use_bvs && a.budget == b.budget &&
           (a.budget != RF_BUDGET_REFERENCE || a.n_samples == b.n_samples) && a.seed == b.seed &&
           a.rays_per_thread == b.rays_per_thread && a.block_size == b.block_size && a.per_test == b.per_test &&
           a.hierarchy == b.hierarchy &&
           (!fast || (a.refine == b.refine && a.transform == b.transform &&
                      (!a.refine || margin(a) == margin(b))));
}

extern "C" int32_t rf_opts_same_tallies(const rf_solve_opts *a, const rf_solve_opts *b)
{
    if (!a || !b) return RF_ERR_INVALID_ARG;
    return same_tallies(*a, *b) ? 1 : 0;
}

extern "C" int32_t rf_debug_fail_solve(rf_ctx *ctx, int32_t nth)
{
    if (!ctx || nth < 0) return set_err(ctx, RF_ERR_INVALID_ARG, "bad failure-injection count");
    ctx->fail_in = nth;
    return RF_OK;
}

// Estimated ray-primitive evaluations per ray of the plain scan: one per
// primitive outside the band stacks; per stack one quadric plus about six band
// evaluations when it is searched (guess, check, walk), one per band when it
// is walked.  The fast form's hierarchy costs about as much as a scan of
// 6 sqrt(m) primitives for m primitives outside the stacks (measured on random
// soups of 80 to 2000 primitives, within 10%); the FP64 walk, used where the
// hierarchy holds every primitive, gains as much or more and is weighed the
// same way.  Admission and packing weigh requests with it.
static double scan_tests_per_ray(const rf_scene *s, const rf_solve_opts &o)
{
    if (o.per_test == 1) return (double)s->n;
    double t = 0.0;
    for (const StackInfo &k : s->stack_info) t += 1.0 + (o.per_test == 0 && k.sorted ? 6.0 : (double)k.bands);
    const double m = (double)(s->n - s->stack_prims);
    const bool bvh = !o.use_bvs && o.hierarchy == 0 &&
                     ((o.math == RF_MATH_FAST && s->bvh_on && o.transform != RF_TRANSFORM_MFMA) ||
                      (o.math == RF_MATH_FP64 && s->desc.bvh_on == 2));
    return t + (bvh ? std::min(m, 6.0 * std::sqrt(m)) : m);
}

struct StreamReq {
    rf_scene *scene = nullptr;
    rf_solve_opts opts{};
    double t_submit = 0.0, t_done = 0.0;
    bool done = false;
    std::vector<int64_t> hits, rays;     // accumulated over the request's launches
    int64_t footprint = 0;
    double cost = 0.0;                   // estimated ray-primitive evaluations, whole request
    int32_t next_emitter = 0;            // emitters [0, next_emitter) are solved
};

struct rf_stream {
    rf_ctx *ctx = nullptr;
    int64_t limit = 0;
    int64_t resident = 0;
    int64_t next_id = 1;
    double max_queued = 0.0;             // admission: queued cost limit (0 = none)
    double max_launch = 0.0;             // packing: cost per launch (0 = one launch per option group)
    double queued_cost = 0.0;
    std::map<int64_t, StreamReq> reqs;
    std::vector<int64_t> queued;         // submission order
};

extern "C" int32_t rf_stream_create(rf_ctx *ctx, int64_t limit, rf_stream **out)
{
    if (!ctx || !out || limit <= 0) return set_err(ctx, RF_ERR_INVALID_ARG, "bad stream arguments");
    rf_stream *st = new rf_stream;
    st->ctx = ctx;
    st->limit = limit;
    ctx->streams.insert(st);
    *out = st;
    return RF_OK;
}

extern "C" int32_t rf_stream_set_cost_limits(rf_stream *st, double max_queued_tests, double max_launch_tests)
{
    if (!st) return RF_ERR_INVALID_ARG;
    if (!(max_queued_tests >= 0.0) || !(max_launch_tests >= 0.0))
        return set_err(st->ctx, RF_ERR_INVALID_ARG, "cost limits must be >= 0");
    st->max_queued = max_queued_tests;
    st->max_launch = max_launch_tests;
    return RF_OK;
}

extern "C" double rf_stream_queued_cost(const rf_stream *st) { return st ? st->queued_cost : -1.0; }

extern "C" int32_t rf_stream_destroy(rf_stream *st)
{
    if (!st) return RF_ERR_INVALID_ARG;
    for (auto &kv : st->reqs)
        if (kv.second.scene) rf_scene_destroy(kv.second.scene);
    st->ctx->streams.erase(st);
    delete st;
    return RF_OK;
}

extern "C" int64_t rf_stream_resident_bytes(const rf_stream *st) { return st ? st->resident : -1; }

extern "C" int32_t rf_stream_submit(rf_stream *st, const rf_primitive *prims, int32_t n, const rf_solve_opts *o,
                                    int64_t *id)
{
    if (!st || !id) return RF_ERR_INVALID_ARG;
    rf_ctx *ctx = st->ctx;
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    if (o->use_bvs) return set_err(ctx, RF_ERR_INVALID_ARG, "streaming solves take plain scenes");
    const int64_t fp = rf_scene_footprint(n, 0);
    if (fp < 0) return set_err(ctx, RF_ERR_INVALID_ARG, "bad scene size");
    if (st->resident + fp > st->limit)
        return set_err(ctx, RF_ERR_CAPACITY, "admission: scene footprint exceeds the HBM budget");
    StreamReq r;
    rc = rf_scene_create(ctx, prims, n, nullptr, 0, &r.scene);
    if (rc) return rc;
    std::vector<int64_t> wi;
    rc = emitter_budget(ctx, r.scene, o, wi);
    double rays = 0.0;
    for (int64_t w : wi) rays += 4.0 * (double)w;
    r.cost = rays * scan_tests_per_ray(r.scene, *o);
    // a request costlier than the limit on its own is admitted onto an empty
    // queue only, so it cannot starve
    if (!rc && st->max_queued > 0.0 && !st->queued.empty() && st->queued_cost + r.cost > st->max_queued)
        rc = set_err(ctx, RF_ERR_CAPACITY, "admission: queued cost would exceed the stream's cost limit");
    if (rc) { rf_scene_destroy(r.scene); return rc; }
    r.opts = *o;
    r.footprint = fp;
    r.t_submit = now_ms();
    st->resident += fp;
    st->queued_cost += r.cost;
    const int64_t rid = st->next_id++;
    st->reqs.emplace(rid, std::move(r));
    st->queued.push_back(rid);
    *id = rid;
    return RF_OK;
}

// Cost of the emitters [e0, e1) of a request.
static double req_cost(const StreamReq &r, int32_t e0, int32_t e1)
{
    return r.cost * (double)(e1 - e0) / (double)std::max(1, r.scene->n);
}

// One launch of a step: whole requests, or the next emitter chunk of one
// request costlier than the launch limit.
struct StreamLaunch {
    std::vector<int64_t> ids;
    int32_t e0 = 0, e1 = 0;              // chunk of ids[0] (whole requests: unused)
    bool chunk = false;
};

// The next launch: option groups in order of their first queued request;
// inside the first group the cheapest requests first, packed up to
// max_launch; a request costlier than max_launch on its own goes in chunks of
// whole emitters of about that cost (at least one emitter).
static StreamLaunch plan_next(const rf_stream *st)
{
    StreamLaunch L;
    const rf_solve_opts &o0 = st->reqs.at(st->queued[0]).opts;
    std::vector<int64_t> g;
    for (int64_t rid : st->queued)
        if (same_tallies(st->reqs.at(rid).opts, o0)) g.push_back(rid);
    const auto left = [&](int64_t rid) {
        const StreamReq &r = st->reqs.at(rid);
        return req_cost(r, r.next_emitter, r.scene->n);
    };
    std::stable_sort(g.begin(), g.end(), [&](int64_t a, int64_t b) { return left(a) < left(b); });
    const double cap = st->max_launch;
    const StreamReq &first = st->reqs.at(g[0]);
    if (cap > 0.0 && left(g[0]) > cap) {
        const double per = req_cost(first, 0, 1);
        const int32_t k = (int32_t)std::max(1.0, std::floor(cap / std::max(per, 1e-30)));
        L.chunk = true;
        L.ids = {g[0]};
        L.e0 = first.next_emitter;
        L.e1 = (int32_t)std::min<int64_t>((int64_t)first.scene->n, (int64_t)L.e0 + k);
        return L;
    }
    double c = 0.0;
    for (int64_t rid : g) {
        const double rc = left(rid);
        if (!L.ids.empty() && cap > 0.0 && c + rc > cap) break;
        L.ids.push_back(rid);
        c += rc;
    }
    return L;
}

// Runs one planned launch and folds its tallies into the requests.
static int stream_run(rf_stream *st, const StreamLaunch &L)
{
    const size_t m = L.ids.size();
    std::vector<rf_scene *> sc;
    std::vector<uint32_t> gids;
    std::vector<int64_t> grays;
    std::vector<std::array<int32_t, 2>> ems;
    std::vector<std::vector<int64_t>> hbuf(m), rbuf(m);
    std::vector<int64_t *> hp, rp;
    for (size_t i = 0; i < m; ++i) {
        StreamReq &r = st->reqs.at(L.ids[i]);
        sc.push_back(r.scene);
        gids.push_back(r.opts.geometry);
        grays.push_back(r.opts.rays_per_emitter);
        ems.push_back(L.chunk ? std::array<int32_t, 2>{L.e0, L.e1}
                              : std::array<int32_t, 2>{r.next_emitter, (int32_t)r.scene->n});
        hbuf[i].resize((size_t)r.scene->n * r.scene->n);
        rbuf[i].resize(r.scene->n);
        hp.push_back(hbuf[i].data());
        rp.push_back(rbuf[i].data());
    }
    rf_solve_opts o = st->reqs.at(L.ids[0]).opts;
    o.form = RF_FORM_PERSISTENT;
    static_assert(sizeof(std::array<int32_t, 2>) == 2 * sizeof(int32_t), "emitter ranges");
    const int rc = solve_impl(st->ctx, sc.data(), (int)m, &o, hp.data(), rp.data(), nullptr, gids.data(),
                              grays.data(), reinterpret_cast<const int32_t(*)[2]>(ems.data()));
    if (rc) return rc;
    const double t = now_ms();
    for (size_t i = 0; i < m; ++i) {
        StreamReq &r = st->reqs.at(L.ids[i]);
        const int32_t e0 = ems[i][0], e1 = ems[i][1], n = r.scene->n;
        if (r.hits.empty()) {
            r.hits.assign((size_t)n * n, 0);
            r.rays.assign(n, 0);
        }
        // the launch solved rows [e0, e1); the others are zero
        for (int32_t e = e0; e < e1; ++e) {
            std::memcpy(&r.hits[(size_t)e * n], &hbuf[i][(size_t)e * n], (size_t)n * sizeof(int64_t));
            r.rays[e] = rbuf[i][e];
        }
        st->queued_cost -= req_cost(r, e0, e1);
        r.next_emitter = e1;
        if (e1 >= n) {
            r.done = true;
            r.t_done = t;
        }
    }
    return RF_OK;
}

extern "C" int32_t rf_stream_step(rf_stream *st, int32_t *n_completed)
{
    if (!st) return RF_ERR_INVALID_ARG;
    if (n_completed) *n_completed = 0;
    if (st->queued.empty()) return RF_OK;
    const StreamLaunch L = plan_next(st);
    const int rc = stream_run(st, L);   // on an error nothing changed: the requests stay queued
    if (rc) return rc;
    std::vector<int64_t> left;
    int done = 0;
    for (int64_t rid : st->queued) {
        if (st->reqs.at(rid).done) ++done;
        else left.push_back(rid);
    }
    st->queued.swap(left);
    if (st->queued.empty()) st->queued_cost = 0.0;   // rounding of the per-emitter split
    if (n_completed) *n_completed = done;
    return RF_OK;
}

extern "C" int32_t rf_stream_flush(rf_stream *st, int32_t *n_completed)
{
    if (!st) return RF_ERR_INVALID_ARG;
    if (n_completed) *n_completed = 0;
    int total = 0;
    while (!st->queued.empty()) {
        int32_t k = 0;
        const int rc = rf_stream_step(st, &k);
        total += k;
        if (n_completed) *n_completed = total;
        if (rc) return rc;
    }
    return RF_OK;
}

extern "C" int32_t rf_stream_collect(rf_stream *st, int64_t id, int64_t *hits, int64_t *rays, double *latency_ms)
{
    if (!st) return RF_ERR_INVALID_ARG;
    auto it = st->reqs.find(id);
    if (it == st->reqs.end() || !it->second.done) return set_err(st->ctx, RF_ERR_INVALID_ARG, "request not completed");
    StreamReq &r = it->second;
    if (hits) std::memcpy(hits, r.hits.data(), r.hits.size() * sizeof(int64_t));
    if (rays) std::memcpy(rays, r.rays.data(), r.rays.size() * sizeof(int64_t));
    if (latency_ms) *latency_ms = r.t_done - r.t_submit;
    st->resident -= r.footprint;
    rf_scene_destroy(r.scene);
    st->reqs.erase(it);
    return RF_OK;
}
