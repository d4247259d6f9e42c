// rf_internal.hpp -- structures shared by the kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rayfactor.h"
#include "rf_device.hpp"

namespace rf {

// Fast-path type segments, in traversal order.
enum SegKind { SEG_CYL = 0, SEG_FRUSTUM = 1, SEG_SPHERE = 2, SEG_RING = 3, SEG_RECT = 4, N_SEG = 5 };

// Per-BV metadata of the fast / fp64 BV kernels.
struct BvMeta {
    int32_t settings;
    int32_t start;
    int32_t end;
    int32_t type;
};

// Runs of the fast form's cylinder segment, in traversal order.  A stack
// (stack >= 1) is a run of two or more consecutive cylinders whose folded
// transform rows are bitwise equal except the z-row offset r[11] -- the bands
// of one coaxial wall: every band shares the ray's x-y quadric and its z
// direction, so the kernel solves the quadric once per ray and tests each
// band's z range.  stack = 2 (3): the offsets are finite and non-decreasing
// (non-increasing) over the run, so a root's bands can be searched for; inv
// is then (end - begin - 1) / (last offset - first offset), 0 if they are
// equal (the search's first guess assumes even spacing).  Other runs
// (stack = 0) are scanned one test at a time.
//
// Analytic band positions (fast form; an_margin > 0 enables them): a searched
// stack with no x or y term in its z row whose offsets lie within an_margin's
// deviation part of the progression r0 + k*delta (k = position - begin,
// |delta| >= 1, so the bands do not overlap).  The root at z offset Z from
// band `begin` then lies at position x = -Z/delta = Z * an_ninv (evaluated in
// FP64), and a position is taken without reading any band record when x is
// far enough from a band boundary that the per-band FP32 tests cannot decide
// otherwise (margin: tau + an_margin + 2^-23 (|c_z t| + 4.1) in z units).
// an_ainv = |an_ninv| in FP32; the caller index at position p is
// orig0 + (p - begin) * ostep, or the band record's when ostep = 0.
struct CylRun {
    int32_t begin;
    int32_t end;
    int32_t stack;
    float inv;
    float r0;
    float an_margin;
    float an_ainv;
    int32_t orig0;
    double an_ninv;
    int32_t ostep;
    int32_t pad;
};
static_assert(sizeof(CylRun) == 48, "CylRun layout");

// Node of the bounding-volume hierarchy over the fast form's per-test
// primitives (GeomDesc::bvh_root), 32 bytes: the boxes of its two children
// as IEEE half-precision bit patterns in the scene's box frame (x - bvh_c) /
// bvh_s (bvh_s a power of two), lower corners rounded down and upper corners
// up, so each holds the padded world box; then one word per child k:
// ref[k] = index | count << 26 | segment << 29 with index sign-extended from
// 26 bits.  count > 0: a leaf of the count records bvh_prims[index ..] of
// type segment; count == 0, index >= 0: the node index; index == -1: no
// child (its box is never tested).
struct BvhNode {
    uint16_t box[12];     // lo0 xyz, hi0 xyz, lo1 xyz, hi1 xyz
    uint32_t ref[2];
};
static_assert(sizeof(BvhNode) == 32, "BvhNode layout");
constexpr int BVH_INDEX_BITS = 26;
constexpr uint32_t BVH_NO_CHILD = (1u << BVH_INDEX_BITS) - 1u;

// Deepest hierarchy a scene may have: a walk keeps at most one stack entry
// per level, and a launch gives each lane as many LDS entries as its deepest
// scene needs.
constexpr int BVH_MAX_DEPTH = 40;

// Per-position band record of a stack: the z-row offset and the primitive's
// index in the caller's order.
struct BandRec {
    float r11;
    int32_t orig;
};

// Device-side descriptor of one geometry (all pointers are device memory).
struct GeomDesc {
    const rf_primitive *prims;          // caller's order (AoS, 112 B)
    const rfd::FastPrim *fast;          // type-segmented traversal order
    const rfd::FastPrim *fast_orig;     // caller's order (BV child scans)
    const rf_bounding_volume *bvs;      // caller's BVs (AoS, 72 B)
    const rfd::FastPrim *fast_bvs;      // BV hit records
    const BvMeta *bv_meta;
    unsigned long long *hits;           // n x n tally, row = emitter
    int32_t n;
    int32_t n_bvs;
    int32_t seg_begin[N_SEG];
    int32_t seg_end[N_SEG];
    uint32_t geometry_id;               // seeded RNG geometry word
    int32_t zring_end;                  // fast: ring positions [seg_begin[SEG_RING], zring_end) are
                                        // world z-plane rings (zring_loop's record encoding)
    // MFMA form: primitive quads (64 floats each: the 16x4 A operand of
    // v_mfma_f32_16x16x4f32, element A[4p + row][k] at a[16k + 4p + row]),
    // type segments padded to whole quads, and traversal position -> index.
    const float *quads;
    const int32_t *pos2orig;
    const int32_t *orig2pos;
    int32_t qseg_begin[N_SEG];
    int32_t qseg_end[N_SEG];
    // Children of each bounding surface in type-segmented order (binned BV
    // kernel): segment s of BV b spans [bvc_seg[b*(N_SEG+1)+s], bvc_seg[b*(N_SEG+1)+s+1])
    // of bvc_fast, whose records carry the child's index in the caller's order.
    const rfd::FastPrim *bvc_fast;
    const int32_t *bvc_seg;
    int32_t bv_binned;                  // 1: no BV asks for the intermediate check (settings != 1)
    int32_t n_cyl_runs;
    const CylRun *cyl_runs;             // cover [seg_begin[SEG_CYL], seg_end[SEG_CYL]) in order
    const BandRec *bands;               // indexed by traversal position (cylinder positions only);
                                        // bands[-1] is a pad record
    // binned BV form with band stacks: runs of the children's cylinder
    // segment of surface b are bvc_runs[bvc_run_off[b] .. bvc_run_off[b+1]),
    // over bvc_fast positions; bvc_bands as bands; bv_cflops[4b + s] the
    // children's algorithmic flops per ray, per test (s = 0) or, with band
    // stacks (s = 1), those of the children outside the stacks;
    // bv_cflops[4b + 2 + s] the number of those children
    const CylRun *bvc_runs;
    const int32_t *bvc_run_off;
    const BandRec *bvc_bands;
    const uint32_t *bv_cflops;
    // compat scan with band stacks: runs over the caller's order (stacks of
    // consecutive cylinders whose invm differs only in invm[11]), band records
    // {invm[11], index} by caller position after one pad record; c_run_of
    // locates the run of a bounding surface's first child
    const CylRun *c_runs;
    const BandRec *c_bands;
    const int32_t *c_run_of;            // run index of each caller position
    int32_t n_c_runs;
    int32_t pad3;
    // the caller's primitives with their emission frames (fast form)
    const rfd::Emitter *emitters;
    // fast form, scenes with many primitives outside band stacks: a
    // hierarchy over those primitives (root node bvh_root), traversed instead
    // of their segment scans, its leaves' records copied to bvh_prims in leaf
    // order with FastPrim::type holding the record's traversal position;
    // bvh_on = 0: no hierarchy (every ring may then be a world z-plane ring)
    const BvhNode *bvh_nodes;
    const rfd::FastPrim *bvh_prims;
    int32_t bvh_root;
    int32_t bvh_on;
    // the nodes' box frame: world x = bvh_c + bvh_s * (stored half value)
    float bvh_c[3];
    float bvh_s;
};

// One work unit: work items [wi0, wi0+n_wi) of one emitter of one geometry.
struct WorkUnit {
    int32_t geom;
    int32_t emitter;
    int64_t wi0;
    int64_t n_wi;
    int64_t pad;
};

struct KParams {
    const GeomDesc *geoms;
    const WorkUnit *units;              // nullptr: implicit unit list (below)
    unsigned int *queue_head;           // atomic unit counter (zeroed before launch)
    unsigned long long *stats;          // [0] tests executed (BV kernels), [1] FP64 re-solves,
                                        // [2] re-scans, [3] flops counted on the device (band
                                        // stacks; the binned BV kernel's child scans), [4] band-
                                        // stack evaluations of the plain scans (kNStats entries)
    int32_t *trace_out;                 // per-ray hit index: ray 4*(wi - trace_wi0) + v
    int64_t trace_wi0;
    int32_t n_units;
    int32_t rng;
    uint32_t seed;
    int32_t refine;
    int32_t bins;                       // LDS tally bins (0 = global atomics per ray)
    // Implicit unit list (units == nullptr): unit u is work items
    // [imp_wi0 + u*imp_unit, min(imp_wi0 + (u+1)*imp_unit, imp_wi0 + imp_n_wi))
    // of emitter imp_emitter of geoms[0], seeded with geometry word imp_geom_id;
    // with imp_emitter < 0, unit u is chunk u mod imp_upe of emitter u / imp_upe
    // (every emitter with the budget imp_n_wi).
    int32_t imp_emitter;
    uint32_t imp_geom_id;
    int64_t imp_wi0;
    int64_t imp_n_wi;
    int64_t imp_unit;
    int32_t imp_upe;
    int32_t pad2;
    unsigned long long *hits;           // tally into this n x n matrix instead of geoms[0]'s (nullptr: the geometry's)
    float tau;                          // fast form: relative margin of the FP64 re-solve flags
    int32_t bv_slots;                   // binned BV kernel: LDS buckets (>= n_bvs of every geometry)
    int32_t stacks;                     // plain (non-BV) forms: 1 = band stacks share the quadric (CylRun)
};

enum KMath { KM_COMPAT = 0, KM_FP64 = 1, KM_FAST = 2 };

constexpr int kNStats = 5;              // KParams::stats entries

struct LaunchCfg {
    int math;        // KMath
    int use_bv;
    int trace;
    int nr;          // rays per thread (fast plain): 1, 2, 4, 8
    int refine;
    int block;       // threads per workgroup
    int grid;        // workgroups
    int direct;      // 1: reference form (one work item per thread, global atomics, no queue)
    int mfma;        // 1: fast plain form with the transform on the matrix cores
    int bvbin;       // 1: fast BV form with rays binned by bounding surface
    int count;       // fast band search: 1 = count its evaluations on the device
    int bt;          // fast band search: 1 = band records copied to LDS per unit
    int bvh;         // fast plain form: units of scenes with a hierarchy walk it (run_unit_bvh)
    size_t lds_bytes;
};

hipError_t launch_solve(const LaunchCfg &cfg, const KParams &kp, hipStream_t stream);
hipError_t launch_fill_u32(unsigned int *p, unsigned int value, size_t count, hipStream_t stream);
hipError_t launch_fill_u64(unsigned long long *p, size_t count, hipStream_t stream);
hipError_t launch_fill_solve(const GeomDesc *geoms, int n_geoms, size_t cells, unsigned long long *stats,
                             unsigned int *queue_head, hipStream_t stream);
hipError_t launch_normalise(const unsigned long long *hits, double *vf, size_t n2, double rays, hipStream_t stream);
hipError_t launch_selftest_math(unsigned long long *counts, unsigned int *first, hipStream_t stream);
hipError_t launch_selftest_sincos(unsigned int *err, hipStream_t stream);

// Algorithmic FP32 flops of one band stack, same counting as flops_per_test,
// counted on the device from the band loop a wave runs: per (ray, stack) the
// x-y rows of origin and direction, the z direction, A, b, the cross term, D,
// sqrt, rcp and both roots; per (ray, band) the band's z origin row (one FMA
// when the z row has no x or y term) and z at each root the wave evaluates
// (root 1 is skipped when no ray of the wave has it ahead).
constexpr uint32_t kFlopsStack = 45;
constexpr uint32_t kFlopsStackC = 48;     // compat: cylinderHit's own A, B, C and roots
constexpr uint32_t kFlopsBandZ0 = 6;
constexpr uint32_t kFlopsBandZ0Axial = 2;
constexpr uint32_t kFlopsBandRoot = 2;
// FP64 form: per (ray, stack) the shared rows, A, B, C, D, sqrt, q, both
// roots and c_z t at each; per (ray, band) z origin and z at both roots.
constexpr uint32_t kFlopsStackD = 51;
constexpr uint32_t kFlopsBandD = 3;

// World z-plane ring of the fast scan (zring_loop): t (subtract, multiply),
// the hit's offsets from the centre (an FMA and a subtract each) and r^2 (a
// multiply and an FMA); once per ray, the reciprocal of c_z.
constexpr int64_t kFlopsZRing = 11;
constexpr int64_t kFlopsZRingRay = 1;

// Hierarchy traversal: the slab test of one child box (six FMAs; the
// min/max reductions are compares) and, once per ray, the three reciprocals
// of the direction and the origin terms (three multiplies).
constexpr uint32_t kFlopsBox = 12;
constexpr uint32_t kFlopsBoxRay = 6;

// Algorithmic FP32 flops per ray-primitive test (counted from the reference
// formulas: FMA = 2, mul/add = 1, rcp/sqrt = 1, compare/select = 0).
__host__ __device__ constexpr int64_t flops_per_test(int type)
{
    switch (type) {
    case RF_RECTANGLE: return 39;
    case RF_DISC: return 42;
    case RF_ANNULUS: return 42;
    case RF_CYLINDER: return 56;
    case RF_FRUSTUM: return 65;
    case RF_SPHERE: return 58;
    default: return 0;
    }
}

}  // namespace rf
