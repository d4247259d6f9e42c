/*
 * rayfactor.h -- C ABI of the MI355X (gfx950) Monte Carlo view-factor engine.
 *
 * Drop-in boundary: this ABI replaces the OpenCL kernel dispatch inside
 * RayFactorCL's runCL_GPURNG (RayFactorCL/main.cpp:54-251).  Plain C types,
 * pointers and sizes only; every failure is returned as an rf_status code
 * (the reference prints and calls exit(1), util_opencl.h:123-131).
 *
 * Threading: an rf_ctx owns one HIP device and one stream; calls on one
 * context must not overlap (no internal locking).  Different contexts are
 * independent.
 */
#ifndef RAYFACTOR_H
#define RAYFACTOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_ABI_VERSION 3

/* ------------------------------------------------------------------ */
/* Scene ABI                                                           */
/* ------------------------------------------------------------------ */

/* Replaces: enum PrimitiveType, RayfactorConstants.h:25-33 (same values). */
typedef enum rf_primitive_type {
    RF_CYLINDER = 0,
    RF_FRUSTUM = 1,
    RF_SPHERE = 2,
    RF_ANNULUS = 3,
    RF_RECTANGLE = 4,
    RF_DISC = 5,
    RF_NONE = 6
} rf_primitive_type;

/* Replaces: struct Primitive, RayfactorConstants.h:38-45 and
 * RayFactorGPU.cl:37-44.  Byte-identical, 112 bytes:
 *   m     object->world transform, 3x4 row-major
 *   invm  world->object transform, 3x4 row-major
 *   util1 frustum top-radius ratio / annulus inner radius squared
 *   util2 annulus outer radius squared / +-1 normal flip (cylinder, frustum, sphere) */
typedef struct rf_primitive {
    float m[12];
    float invm[12];
    int32_t type;
    int32_t rayDensity;
    float util1;
    float util2;
} rf_primitive;

/* Replaces: struct BoundingVolume, RayfactorConstants.h:48-56 and
 * RayFactorGPU.cl:46-54.  Byte-identical, 72 bytes.  Children are the
 * primitives [objStartIndex, objEndIndex); settings == 1 tests them as soon
 * as the bounding surface is hit ("intermediateCheck"). */
typedef struct rf_bounding_volume {
    float invm[12];
    int32_t type;
    int32_t settings;
    int32_t objStartIndex;
    int32_t objEndIndex;
    float util1;
    float util2;
} rf_bounding_volume;

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
typedef enum rf_status {
    RF_OK = 0,
    RF_ERR_INVALID_ARG = -1,
    RF_ERR_INVALID_SCENE = -2,   /* unknown type, bad BV child range, n == 0 */
    RF_ERR_HIP = -3,             /* HIP runtime error (see rf_last_error) */
    RF_ERR_OUT_OF_MEMORY = -4,
    RF_ERR_OVERFLOW = -5,        /* int32 overflow of a reference quantity */
    RF_ERR_IO = -6,              /* file missing / unwritable */
    RF_ERR_PARSE = -7,           /* malformed scene XML */
    RF_ERR_CAPACITY = -8,        /* request exceeds a configured limit */
    RF_ERR_NO_DEVICE = -9
} rf_status;

/* ------------------------------------------------------------------ */
/* Solve options                                                       */
/* ------------------------------------------------------------------ */
typedef enum rf_math {
    RF_MATH_COMPAT = 0,   /* FP32, bit-exact reference semantics (pinned IEEE builtins) */
    RF_MATH_FP64 = 1,     /* FP64 physics semantics: the error oracle form */
    RF_MATH_FAST = 2      /* FP32 hardware rcp/sqrt, FP64 re-solve of ill-conditioned roots */
} rf_math;

typedef enum rf_rng {
    RF_RNG_REFERENCE = 0, /* key {gid, 0xdecafbad}, ctr {1, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d} */
    RF_RNG_SEEDED = 1     /* key {wi_lo, seed}, ctr {wi_hi, emitter, geometry, 0xbeeff00d} */
} rf_rng;

typedef enum rf_budget {
    RF_BUDGET_REFERENCE = 0, /* main.cpp:172-173: every emitter uses primitive 0's rayDensity */
    RF_BUDGET_EXPLICIT = 1   /* rays_per_emitter (rounded up to whole 4-ray work items) */
} rf_budget;

typedef enum rf_form {
    RF_FORM_PER_EMITTER = 0, /* one launch per emitter (the reference's dispatch shape) */
    RF_FORM_PERSISTENT = 1,  /* one persistent launch fed by a device work queue */
    RF_FORM_REFERENCE = 2    /* one launch per emitter, 128-wide groups, one global atomic per ray */
} rf_form;

typedef struct rf_solve_opts {
    int32_t math;             /* rf_math */
    int32_t rng;              /* rf_rng */
    int32_t use_bvs;          /* 0: runGPU semantics, 1: runBVGPU semantics */
    int32_t budget;           /* rf_budget */
    int32_t n_samples;        /* RF_BUDGET_REFERENCE: the reference's n_samples */
    int32_t form;             /* rf_form */
    int64_t rays_per_emitter; /* RF_BUDGET_EXPLICIT */
    uint32_t seed;            /* RF_RNG_SEEDED */
    uint32_t geometry;        /* RF_RNG_SEEDED: geometry id (batched solves use their index) */
    int32_t refine;           /* RF_MATH_FAST: 1 = FP64 re-solve of flagged roots */
    int32_t rays_per_thread;  /* 0 = autotuned, else the default: 2 when per_test = 0 and a scene has a
                                 sorted band stack, else 4; or 1, 2, 4, 8 */
    int32_t block_size;       /* 0 = autotuned/default; else 64..256, multiple of 64 */
    int32_t transform;        /* RF_MATH_FAST plain form: rf_transform */
    float refine_margin;      /* RF_MATH_FAST: relative margin of the FP64 re-solve flags, 0 = 2^-14;
                                 larger re-solves more pairs (at most 0.25) */
    int32_t per_test;         /* Scans without bounding surfaces (and the fast form's binned BV scan):
                                 0 = a band stack (consecutive cylinders whose transforms differ only
                                 in the z offset, e.g. the bands of one wall) solves its quadric once
                                 per ray; a stack whose z offsets are sorted is then searched for the
                                 bands a root can lie on, any other has each band's z range tested;
                                 1 = every cylinder test in full; 2 = band stacks, each band's z range
                                 tested.  Same tallies in all three, in every arithmetic form. */
    int32_t count_tests;      /* 1 (the default): the fast form's band search counts its evaluations
                                 and flops on the device; 0: it does not (about 2% faster), and
                                 rf_timing.tests and .flops are -1 for a solve that searched stacks.
                                 A launch counted once gives the counts of every identical launch. */
    int32_t band_tile;        /* RF_MATH_FAST band search (1 or 2 rays per thread): each work unit's
                                 band records (its geometry's cylinder segment) are copied to LDS and
                                 searched there when every scene's segment has at most this many
                                 records; 0 = the autotuner's choice (off until tuned; up to 8192
                                 records when on), -1 = off.  Same tallies either way. */
    int32_t hierarchy;        /* Plain scans (no bounding surfaces, per_test 0 or 2): 0 (the default)
                                 = in scenes with at least 48 primitives outside band stacks, the
                                 RF_MATH_FAST form with the VALU transform finds those through a
                                 bounding-box hierarchy (only boxes the ray crosses before its best
                                 hit are opened; each group's rays sorted by direction first; 2 rays
                                 per thread unless rays_per_thread asks), and RF_MATH_FP64 does the
                                 same where the scene has no band stacks; -1 = every primitive
                                 tested.  Same tallies either way, exact ties between primitives
                                 included. */
} rf_solve_opts;

typedef enum rf_transform {
    RF_TRANSFORM_AUTO = 0,    /* autotuned choice (VALU until tuned) */
    RF_TRANSFORM_VALU = 1,    /* ray -> object transform on the vector ALUs */
    RF_TRANSFORM_MFMA = 2     /* ray -> object transform on the matrix cores (v_mfma_f32_16x16x4f32) */
} rf_transform;

/* Fills *opts with defaults: COMPAT math, reference RNG and budget,
 * n_samples = 1000000, persistent form (same tallies as one launch per
 * emitter, one launch for all emitters), refine = 1, count_tests = 1, band_tile = 0. */
void rf_solve_opts_default(rf_solve_opts *opts);

/* Measurements of the last solve on a context. */
typedef struct rf_timing {
    double kernel_ms;         /* HIP events around the device work, on the library stream */
    double wall_ms;           /* host steady clock around the whole call */
    int64_t rays;             /* rays traced */
    int64_t tests;            /* ray-primitive evaluations executed: one per primitive tested in full;
                                 in a band stack one per (ray, stack) for the shared quadric and one per
                                 band whose z range is checked (counted on the device), as in the
                                 bounding-surface forms' child scans; FP64 re-solves are not counted */
    int64_t flops;            /* algorithmic FP32 flops: sum over tests of F_type; a stacked band
                                 test counts F_band, plus F_stack per (ray, band stack) */
    int64_t refined;          /* (ray, primitive) pairs re-solved in FP64 */
    int64_t launches;         /* kernel launches issued */
    int64_t rescans;          /* lane re-scans after a full FP64 re-solve list (no flagged pair is left unsolved) */
    int64_t graph_launches;   /* launches issued as a HIP graph (persistent solves repeated with the
                                 same scenes and options: captured on the second call, replayed after) */
    int64_t tests_brute_force; /* rays x primitives: the tests of runGPU's brute-force scan
                                  (RayFactorGPU.cl:445-472) for the same rays */
    int32_t shape_rays_per_thread; /* launch shape of the last launch: rays per thread (0: MFMA transform) */
    int32_t shape_block;      /* threads per workgroup */
    int32_t shape_flags;      /* RF_SHAPE_* bits */
    int32_t shape_pad;
} rf_timing;

/* rf_timing.shape_flags bits. */
enum {
    RF_SHAPE_MFMA = 1,        /* ray transform on the matrix cores */
    RF_SHAPE_STACKS = 2,      /* band stacks share their quadric */
    RF_SHAPE_SEARCHED = 4,    /* sorted band stacks searched (fast form, 1-2 rays per thread) */
    RF_SHAPE_BAND_TILE = 8,   /* band records read from an LDS copy */
    RF_SHAPE_COUNTED = 16,    /* evaluations and flops counted (always, except an uncounted band search) */
    RF_SHAPE_BV_BINNED = 32,  /* fast bounding-surface form with rays binned by surface */
    RF_SHAPE_REFERENCE = 64,  /* the reference dispatch shape (one work item per thread) */
    RF_SHAPE_HIERARCHY = 128  /* per-test primitives found through bounding-box hierarchies */
};

/* ------------------------------------------------------------------ */
/* Library / context                                                   */
/* ------------------------------------------------------------------ */
typedef struct rf_ctx rf_ctx;
typedef struct rf_scene rf_scene;

int32_t rf_abi_version(void);
const char *rf_status_string(int32_t status);

/* Replaces: opencl_init (util_opencl.h:133-334) -- one HIP device, one stream. */
int32_t rf_ctx_create(int32_t hip_device, rf_ctx **out);
/* Also releases the context's remaining scenes and streams; handles to them
 * are invalid afterwards. */
int32_t rf_ctx_destroy(rf_ctx *ctx);
/* Last error text of a context (empty if none).  Valid until the next call. */
const char *rf_last_error(const rf_ctx *ctx);
/* Device name and compute-unit count (for run reports). */
int32_t rf_device_info(const rf_ctx *ctx, char *name, size_t name_len, int32_t *n_cu, int32_t *clock_mhz);

/* Replaces: clCreateBuffer + blocking clEnqueueWriteBuffer of the objects /
 * bvs / zeroed hits buffers (main.cpp:118-134).  The library copies the
 * arrays to device memory it owns; the caller's arrays are not retained. */
int32_t rf_scene_create(rf_ctx *ctx, const rf_primitive *prims, int32_t n_prims,
                        const rf_bounding_volume *bvs, int32_t n_bvs, rf_scene **out);
int32_t rf_scene_destroy(rf_scene *scene);
int32_t rf_scene_size(const rf_scene *scene, int32_t *n_prims, int32_t *n_bvs);
/* Device bytes held by a scene (primitive layouts plus its n x n int64 tally). */
int64_t rf_scene_footprint(int32_t n_prims, int32_t n_bvs);

/* ------------------------------------------------------------------ */
/* Dispatch                                                            */
/* ------------------------------------------------------------------ */

/* Replaces: main.cpp:172-173, the work items launched per emitter:
 * round(rayDensity*n_samples/count) to a multiple of wgsize, ties down.
 * RF_ERR_OVERFLOW when rayDensity*n_samples overflows int32. */
int32_t rf_nthreads(int32_t ray_density, int32_t n_samples, int32_t count, int32_t wgsize,
                    int64_t *nthreads);

/* Replaces: clSetKernelArg(kernel, 0, &objNo) + clEnqueueNDRangeKernel
 * (main.cpp:170-176) of runGPU (RayFactorGPU.cl:481) or runBVGPU (:530):
 * work items [wi0, wi0 + n_wi) of one emitter, 4 rays each, tallied into the
 * scene's device hit matrix.  Asynchronous on the context stream. */
int32_t rf_launch_emitter(rf_ctx *ctx, rf_scene *scene, int32_t emitter, int64_t wi0,
                          int64_t n_wi, const rf_solve_opts *opts);
/* Replaces: clFinish (main.cpp:185-188). */
int32_t rf_sync(rf_ctx *ctx);
/* Replaces: the blocking clEnqueueReadBuffer of hits (main.cpp:199-205); int64 cells. */
int32_t rf_hits_read(rf_ctx *ctx, rf_scene *scene, int64_t *hits);
/* Replaces: the normalisation of main.cpp:216-246 (float division of the
 * summed hits by samplesForObj), done on the device in FP64: vf[i][j] =
 * hits[i][j] / rays_per_emitter for the scene's device tally (every emitter of
 * a solve has the same budget; 0 where it is 0).  Bit-identical to
 * rf_normalise on the same counts. */
int32_t rf_vf_read(rf_ctx *ctx, rf_scene *scene, int64_t rays_per_emitter, double *vf);
/* Replaces: the blocking zero write of hits (main.cpp:130-134). */
int32_t rf_hits_reset(rf_ctx *ctx, rf_scene *scene);

/* Replaces: the whole dispatch of runCL_GPURNG (main.cpp:54-214): every
 * emitter, the launch geometry, and the readback.  hits: caller-owned n*n
 * int64 (row = emitter); rays: caller-owned n (rays traced per emitter);
 * timing may be NULL. */
int32_t rf_solve(rf_ctx *ctx, rf_scene *scene, const rf_solve_opts *opts,
                 int64_t *hits, int64_t *rays, rf_timing *timing);

/* The persistent form of rf_solve for callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation.  No counterpart in the reference.  All emitters of the scene,
 * in one launch on the context stream, tallied into the caller's device
 * matrix d_hits (n*n uint64, row = emitter, cleared by the call); the same
 * tallies as rf_solve with the same options (opts->form must be
 * RF_FORM_PERSISTENT).  Asynchronous: rf_sync, then read d_hits.  It allocates
 * nothing: the queue counter and statistics are a fixed 256-byte device array
 * per GPU, held until the process ends and shared by every call on that GPU,
 * so calls that overlap on one GPU (from any context or host thread) race.
 * Limit: 2^31-1 work units (RF_ERR_CAPACITY beyond, before any launch); no
 * timing is reported. */
int32_t rf_solve_noalloc(rf_ctx *ctx, rf_scene *scene, const rf_solve_opts *opts, uint64_t *d_hits);

/* Per-ray hit index (-1 = miss) of work items [wi0, wi0+n_wi) of one
 * emitter: idx_out holds 4*n_wi int32.  No reference counterpart: the
 * diagnostic behind the differing-ray counts between arithmetic modes. */
int32_t rf_trace_rays(rf_ctx *ctx, rf_scene *scene, int32_t emitter, int64_t wi0, int64_t n_wi,
                      const rf_solve_opts *opts, int32_t *idx_out);

/* Replaces: runCL_GPURNG (main.cpp:54-272) with its exact observables:
 * int32 hits (RF_ERR_OVERFLOW past 2^31-1 instead of wrapping) and
 * samplesForObj = nthreads*4 (main.cpp:178).  vfm (may be NULL) receives the
 * reference's float normalisation (main.cpp:216-246). */
int32_t rf_run_cl_gpurng(rf_ctx *ctx, const rf_primitive *prims, int32_t n_prims,
                         const rf_bounding_volume *bvs, int32_t n_bvs, int32_t use_bvs,
                         int32_t n_samples, int32_t *hits, double *samples_for_obj, float *vfm);

/* ------------------------------------------------------------------ */
/* Batched sweeps (no reference counterpart: furnace design sweeps)    */
/* ------------------------------------------------------------------ */

/* Solves n_geoms independent scenes.  RF_FORM_PERSISTENT runs all their
 * (geometry, emitter, ray-chunk) work units in one persistent launch fed by
 * a device work queue (a HIP graph from the second identical call on); RF_FORM_PER_EMITTER /
 * RF_FORM_REFERENCE issue one launch per emitter of each geometry.
 * hits[g] is a caller-owned n_g*n_g int64 matrix (may be NULL to skip the
 * readback); rays[g] (may be NULL) receives n_g int64.  With RF_RNG_SEEDED,
 * geometry g uses geometry id opts->geometry + g. */
int32_t rf_batch_solve(rf_ctx *ctx, rf_scene *const *scenes, int32_t n_geoms,
                       const rf_solve_opts *opts, int64_t *const *hits, int64_t *const *rays,
                       rf_timing *timing);

/* Autotuner: times the FAST kernel over {VALU transform x rays per thread
 * (2 with and without the LDS band tile, 4, 8), MFMA transform} x {block size}
 * on the given scenes (a bounded sample of their work) and stores the fastest
 * choice in the context, where later solves with rays_per_thread == 0,
 * block_size == 0, transform == AUTO and band_tile == 0 use it.  best_rpt
 * (0 for the MFMA transform) / best_block may be NULL. */
int32_t rf_autotune(rf_ctx *ctx, rf_scene *const *scenes, int32_t n_geoms,
                    const rf_solve_opts *opts, int32_t *best_rpt, int32_t *best_block);

/* ------------------------------------------------------------------ */
/* Streaming solves (resident scenes, HBM-footprint admission)         */
/* ------------------------------------------------------------------ */
typedef struct rf_stream rf_stream;

/* A service that keeps scenes and tallies resident on the device and batches
 * queued requests of unequal size into persistent launches.  Admission
 * rejects (RF_ERR_CAPACITY) a request whose scene footprint would push the
 * resident total past hbm_limit_bytes.  No reference counterpart (the
 * reference solves one scene per process run, main.cpp:275-284). */
int32_t rf_stream_create(rf_ctx *ctx, int64_t hbm_limit_bytes, rf_stream **out);
int32_t rf_stream_destroy(rf_stream *st);
/* Cost limits, in estimated ray-primitive evaluations (rays x the plain
 * scan's evaluations per ray: one per primitive outside band stacks, per stack
 * one quadric plus ~6 band evaluations when searched or one per band when
 * walked).  max_queued_tests: admission rejects (RF_ERR_CAPACITY) a submit
 * that would push the queued, not yet flushed, cost past it; a request
 * costlier than the limit on its own is admitted only onto an empty queue.
 * max_launch_tests: a flush packs each option group's requests, cheapest
 * first, into persistent launches of at most this cost (a costlier request
 * runs alone).  0 = no limit (the defaults). */
int32_t rf_stream_set_cost_limits(rf_stream *st, double max_queued_tests, double max_launch_tests);
/* Estimated cost of the queued requests (-1 for a NULL stream). */
double rf_stream_queued_cost(const rf_stream *st);
/* Enqueues one solve; the primitives are copied.  *request_id identifies it. */
int32_t rf_stream_submit(rf_stream *st, const rf_primitive *prims, int32_t n_prims,
                         const rf_solve_opts *opts, int64_t *request_id);
/* Runs one persistent launch of queued work and returns: requests share a
 * launch only when rf_opts_same_tallies holds for their options; the first
 * queued request's option group goes first, cheapest requests first, packed up
 * to the launch limit of rf_stream_set_cost_limits; a request costlier than
 * the limit runs in chunks of whole emitters of about that cost, one chunk per
 * step, so requests submitted in between can run before its next chunk.
 * *n_completed counts the requests this step completed.  On an error nothing
 * of the launch is kept and its requests stay queued. */
int32_t rf_stream_step(rf_stream *st, int32_t *n_completed);
/* rf_stream_step until the queue is empty; *n_completed counts the requests
 * completed (also when an error stops it; the rest stay queued). */
int32_t rf_stream_flush(rf_stream *st, int32_t *n_completed);
/* Copies a completed request's n*n hits (and per-emitter rays) out and
 * releases its device memory.  RF_ERR_INVALID_ARG if not completed. */
int32_t rf_stream_collect(rf_stream *st, int64_t request_id, int64_t *hits, int64_t *rays,
                          double *latency_ms);
int64_t rf_stream_resident_bytes(const rf_stream *st);
/* 1 if two solves may share one persistent launch: equal math, RNG, bounding
 * surface use, budget rule (and n_samples under the reference budget), seed,
 * launch shape and scan form, and for the fast form equal FP64 re-solve,
 * re-solve margin and transform unit -- every option that can change a
 * tally.  0 otherwise; RF_ERR_INVALID_ARG for a NULL argument.  Host only. */
int32_t rf_opts_same_tallies(const rf_solve_opts *a, const rf_solve_opts *b);
/* Test hook (no reference counterpart): the nth solve issued on the context
 * from now (1 = the next) fails with RF_ERR_HIP before any device work; 0
 * cancels.  Used to check the streaming service's recovery. */
int32_t rf_debug_fail_solve(rf_ctx *ctx, int32_t nth);

/* ------------------------------------------------------------------ */
/* Host-side scene input and result output                             */
/* ------------------------------------------------------------------ */

/* A scene read from the reference XML format. */
typedef struct rf_host_scene {
    rf_primitive *prims;
    int32_t n_prims;
    rf_bounding_volume *bvs;
    int32_t n_bvs;
    float globalRayDensity;
    int32_t useBoundingVolumes;
} rf_host_scene;

/* Replaces: Scene::readScene (RayFactorCL/Scene.cpp:24-452): settings,
 * primitives, bounding volumes, affine transforms and per-type rayDensity,
 * producing the same bytes (fields the reference leaves uninitialised are
 * zero).  Free with rf_host_scene_free. */
int32_t rf_scene_load_xml(const char *path, rf_host_scene *out);
/* The same from an in-memory XML document. */
int32_t rf_scene_parse_xml(const char *text, size_t len, rf_host_scene *out);
void rf_host_scene_free(rf_host_scene *scene);

/* Replaces: the output.txt writer (main.cpp:236-251): row i is
 * float(hits[i][j]) / samples[i] printed with precision 8 and a trailing
 * comma after every value. */
int32_t rf_write_output(const char *path, int32_t n, const int64_t *hits, const double *samples);
/* FP64 normalisation: vf[i][j] = hits[i][j] / rays[i] (0 where rays[i] == 0). */
int32_t rf_normalise(int32_t n, const int64_t *hits, const int64_t *rays, double *vf);

/* Device self-test of the compat arithmetic (no reference counterpart).  The
 * compat kernels pin native_recip / native_sqrt to IEEE 1/x and sqrt and reach
 * them through short hardware sequences (a Newton step, a one-ulp correction)
 * that report when an operand is outside their exact range; the test is then
 * repeated with the IEEE operations.  This checks, over all 2^32 float inputs,
 * that the sequences give the compiler's IEEE results wherever they report
 * success: mismatches[0] (reciprocal) and mismatches[1] (square root) are 0
 * on a conforming device.  [2] counts the Newton reciprocal's misses on all
 * inputs and [3] the raw v_sqrt_f32's misses on positive normal inputs
 * (information).
 * first[k] = lowest mismatching input bits (0xffffffff if none). */
int32_t rf_selftest_math(rf_ctx *ctx, uint64_t mismatches[4], uint32_t first[4]);

/* Largest absolute error, against FP64, of the fast form's sin and cos of
 * 2 pi u (v_sin_f32 / v_cos_f32, argument in revolutions) over every
 * emission uniform u = k 2^-24, k < 2^24: max_err[0] sin, max_err[1] cos.
 * No reference counterpart (RayFactorGPU.cl:114 calls the vendor's sincos,
 * whose accuracy is the runtime's): a device check of the builtin the fast
 * emission uses. */
int32_t rf_selftest_sincos(rf_ctx *ctx, double max_err[2]);

/* Surface area of each primitive: its object-space shape (unit cylinder
 * z in [0,1], frustum radius 1+(util1-1)z, unit sphere, unit disc, annulus
 * r^2 in [util1, util2], rectangle [-1,1]^2) under the linear part of m,
 * integrated by quadrature (exact for the reference's scalings and
 * rotations).  No reference counterpart: input of the reciprocity check. */
int32_t rf_areas(const rf_primitive *prims, int32_t n, double *area);

/* Closure and reciprocity of a view-factor matrix (row = emitter), the
 * checks beside the output writer.  closure (may be NULL) receives
 * sum_j vf[i][j].  Reciprocity compares A_i F_ij with A_j F_ji over pairs
 * where both are at least min_product; rel is |d| / max of the two. */
typedef struct rf_vf_report {
    double min_closure, max_closure;  /* over emitters */
    double max_recip_abs;             /* max |A_i F_ij - A_j F_ji| */
    double max_recip_rel;             /* max of the same relative to the larger product */
    int32_t worst_i, worst_j;         /* pair of max_recip_rel (-1 if none compared) */
    int32_t pairs_compared;
} rf_vf_report;
int32_t rf_vf_check(int32_t n, const double *vf, const double *area, double min_product, double *closure,
                    rf_vf_report *report);

#ifdef __cplusplus
}
#endif
#endif /* RAYFACTOR_H */
