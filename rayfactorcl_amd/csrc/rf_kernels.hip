// rf_kernels.hip -- hand-written gfx950 kernels of the view-factor engine.
//
// One kernel family replaces RayFactorCL's runGPU / runBVGPU
// (RayFactorGPU.cl:481-626).  A workgroup takes work units -- (geometry,
// emitter, range of 4-ray work items) -- from a device queue, traces its
// rays against the scene, and tallies nearest hits in LDS (one 32-bit bin per
// primitive), flushing non-zero bins to the int64 hit matrix with one global
// atomic each.  Scene records are read through the constant address space
// at wave-uniform indices, so they arrive in SGPRs (s_load) and every lane
// of a wave walks the same primitive at the same time: the type dispatch
// never diverges.
//
//   compat  FP32 reference semantics, caller's primitive order
//   fp64    FP64 physics semantics, caller's primitive order
//   fast    FP32 with hardware rcp/sqrt over type-segmented primitive
//           records, NR rays per thread, FP64 re-solve of flagged roots
//
// "direct" launches keep the reference's dispatch shape instead: one work
// item per thread, 128-wide groups, one global atomic per hitting ray.
#include "rf_internal.hpp"

#pragma clang fp contract(off)

namespace rf {

using namespace rfd;

#define CPTR(T) const __attribute__((address_space(4))) T *

template <typename T>
__device__ __forceinline__ CPTR(T) cst(const T *p) { return (CPTR(T))p; }

// Copy of a record read through the constant address space, dword by dword
// (unused fields are dropped by the compiler; the rest become s_load).
template <typename T>
__device__ __forceinline__ T ldc(CPTR(T) p)
{
    static_assert(sizeof(T) % 4 == 0, "record must be dword sized");
    uint32_t d[sizeof(T) / 4];
    CPTR(uint32_t) q = (CPTR(uint32_t))p;
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) d[i] = q[i];
    T v;
    __builtin_memcpy(&v, d, sizeof(T));
    return v;
}

// The v-th ray of a work item uses rot^(v+1) of the random block
// (rands = rands.yzwx before every ray, RayFactorGPU.cl:502-506).
__device__ __forceinline__ void rotate_rands(const float f[4], int v, float r4[4])
{
    r4[0] = f[(v + 1) & 3];
    r4[1] = f[(v + 2) & 3];
    r4[2] = f[(v + 3) & 3];
    r4[3] = f[v & 3];
}

__device__ __forceinline__ void workitem_floats(const KParams &kp, uint32_t geometry, int emitter,
                                                int64_t wi, float f[4])
{
    const uint4 u = workitem_u32(kp.rng, kp.seed, geometry, emitter, wi);
    f[0] = u01_24(u.x); f[1] = u01_24(u.y); f[2] = u01_24(u.z); f[3] = u01_24(u.w);
}

// ---------------------------------------------------------------------------
// compat / fp64: findObject (RayFactorGPU.cl:445-472) and the BV body
// (:578-624) over the caller's primitive order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_c(CPTR(rf_primitive) P, int start, int end, const Ray32 &ray)
{
    int obj = -1;
    float tbest = finf();
    for (int w = start; w < end; ++w) {
        const rf_primitive p = ldc(P + w);
        const float t = hit_c(p.type, p.invm, p.util1, p.util2, ray);
        if (t < tbest) { tbest = t; obj = w; }
    }
    return obj;
}

__device__ __forceinline__ int find_d(CPTR(rf_primitive) P, int start, int end, const Ray64 &ray)
{
    int obj = -1;
    double tbest = dinf();
    for (int w = start; w < end; ++w) {
        const rf_primitive p = ldc(P + w);
        const double t = hit_d(p.type, p.invm, p.util1, p.util2, ray);
        if (t < tbest) { tbest = t; obj = w; }
    }
    return obj;
}

// BV body shared by compat and fp64 (MATH selects the arithmetic).  The
// final child scan of lanes whose winning bounding surface differs is done
// one distinct surface at a time: the leader's surface is read with
// readlane, so every scan runs over a wave-uniform range.
template <int MATH, typename RayT>
__device__ __forceinline__ int find_bv_cd(CPTR(rf_primitive) P, CPTR(rf_bounding_volume) B, int nbv,
                                          const RayT &ray, uint64_t &tests)
{
    int bvIndex = -1, objectIndex = -1;
    double tbest = dinf();
    for (int w = 0; w < nbv; ++w) {
        const rf_bounding_volume b = ldc(B + w);
        double t;
        if constexpr (MATH == KM_COMPAT) t = hit_c(b.type, b.invm, b.util1, b.util2, ray);
        else t = hit_d(b.type, b.invm, b.util1, b.util2, ray);
        ++tests;
        if (t < tbest) {
            if (b.settings == 1) {
                if constexpr (MATH == KM_COMPAT) objectIndex = find_c(P, b.objStartIndex, b.objEndIndex, ray);
                else objectIndex = find_d(P, b.objStartIndex, b.objEndIndex, ray);
                tests += (uint64_t)(b.objEndIndex - b.objStartIndex);
                if (objectIndex != -1) { tbest = t; bvIndex = w; }
            } else {
                tbest = t; bvIndex = w;
            }
        }
    }
    bool pending = (bvIndex != -1 && objectIndex == -1);
    for (;;) {
        const uint64_t mask = __ballot(pending);
        if (mask == 0) break;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        const int bsel = __builtin_amdgcn_readlane(bvIndex, leader);
        if (pending && bvIndex == bsel) {
            const int s = B[bsel].objStartIndex, e = B[bsel].objEndIndex;
            if constexpr (MATH == KM_COMPAT) objectIndex = find_c(P, s, e, ray);
            else objectIndex = find_d(P, s, e, ray);
            tests += (uint64_t)(e - s);
            pending = false;
        }
    }
    return objectIndex;
}

// ---------------------------------------------------------------------------
// fast: per-type tests over FastPrim records.  Each returns t (+INF on a
// miss) and sets flag when the FP32 roots are ill-conditioned.
// ---------------------------------------------------------------------------
#define ROW_S(P, i, R) __builtin_fmaf(P.r[4 * i], R.sx, __builtin_fmaf(P.r[4 * i + 1], R.sy, __builtin_fmaf(P.r[4 * i + 2], R.sz, P.r[4 * i + 3])))
#define ROW_C(P, i, R) __builtin_fmaf(P.r[4 * i], R.cx, __builtin_fmaf(P.r[4 * i + 1], R.cy, P.r[4 * i + 2] * R.cz))

template <int KIND>
__device__ __forceinline__ float fast_test(const FastPrim &P, const Ray32 &R, bool &flag)
{
    if constexpr (KIND == SEG_CYL) {
        const float sx = ROW_S(P, 0, R), sy = ROW_S(P, 1, R), sz = ROW_S(P, 2, R);
        const float cx = ROW_C(P, 0, R), cy = ROW_C(P, 1, R), cz = ROW_C(P, 2, R);
        const float A = __builtin_fmaf(cx, cx, cy * cy);
        const float b = __builtin_fmaf(sx, cx, sy * cy);
        const float C = __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, -1.0f));
        const float D = __builtin_fmaf(b, b, -(A * C));
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float t1 = (-b - sq) * rA;
        const float t2 = (sq - b) * rA;
        const float z1 = __builtin_fmaf(cz, t1, sz);
        const float z2 = __builtin_fmaf(cz, t2, sz);
        const bool v1 = (t1 > EPS_F) & (__builtin_fabsf(z1) <= 0.5f);
        const bool v2 = (t2 > EPS_F) & (__builtin_fabsf(z2) <= 0.5f);
        flag = (__builtin_fabsf(D) <= TAU_D * (b * b)) |
               (__builtin_fabsf(__builtin_fabsf(z1) - 0.5f) <= TAU_Z) |
               (__builtin_fabsf(__builtin_fabsf(z2) - 0.5f) <= TAU_Z);
        return v1 ? t1 : (v2 ? t2 : finf());
    } else if constexpr (KIND == SEG_FRUSTUM) {
        const float sx = ROW_S(P, 0, R), sy = ROW_S(P, 1, R), sw = ROW_S(P, 2, R);
        const float cx = ROW_C(P, 0, R), cy = ROW_C(P, 1, R), cw = ROW_C(P, 2, R);
        const float Axy = __builtin_fmaf(cx, cx, cy * cy);
        const float A = __builtin_fmaf(-cw, cw, Axy);
        const float b = __builtin_fmaf(sx, cx, __builtin_fmaf(sy, cy, -(sw * cw)));
        const float C = __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, -(sw * sw)));
        const float D = __builtin_fmaf(b, b, -(A * C));
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float ta = (-b - sq) * rA;
        const float tb = (sq - b) * rA;
        const float t1 = __builtin_fminf(ta, tb);
        const float t2 = __builtin_fmaxf(ta, tb);
        const float w1 = __builtin_fmaf(cw, t1, sw) - P.k0;
        const float w2 = __builtin_fmaf(cw, t2, sw) - P.k0;
        const bool v1 = (t1 > EPS_F) & (__builtin_fabsf(w1) <= P.k1);
        const bool v2 = (t2 > EPS_F) & (__builtin_fabsf(w2) <= P.k1);
        const float rim = (2.0f * TAU_Z) * P.k1;
        flag = (__builtin_fabsf(D) <= TAU_D * (b * b)) | (__builtin_fabsf(A) <= TAU_A * Axy) |
               (__builtin_fabsf(__builtin_fabsf(w1) - P.k1) <= rim) |
               (__builtin_fabsf(__builtin_fabsf(w2) - P.k1) <= rim);
        return v1 ? t1 : (v2 ? t2 : finf());
    } else if constexpr (KIND == SEG_SPHERE) {
        const float sx = ROW_S(P, 0, R), sy = ROW_S(P, 1, R), sz = ROW_S(P, 2, R);
        const float cx = ROW_C(P, 0, R), cy = ROW_C(P, 1, R), cz = ROW_C(P, 2, R);
        const float A = __builtin_fmaf(cx, cx, __builtin_fmaf(cy, cy, cz * cz));
        const float b = __builtin_fmaf(sx, cx, __builtin_fmaf(sy, cy, sz * cz));
        const float C = __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, __builtin_fmaf(sz, sz, -1.0f)));
        const float D = __builtin_fmaf(b, b, -(A * C));
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float t1 = (-b - sq) * rA;
        const float t2 = (sq - b) * rA;
        flag = (__builtin_fabsf(D) <= TAU_D * (b * b));
        return (t1 > EPS_F) ? t1 : ((t2 > EPS_F) ? t2 : finf());
    } else {
        // planes: object z = 0; hit point found in world space, then mapped
        const float sz = ROW_S(P, 2, R), cz = ROW_C(P, 2, R);
        const float t = -sz * rcp_fast(cz);
        const float hX = __builtin_fmaf(R.cx, t, R.sx);
        const float hY = __builtin_fmaf(R.cy, t, R.sy);
        const float hZ = __builtin_fmaf(R.cz, t, R.sz);
        const float hx = __builtin_fmaf(P.r[0], hX, __builtin_fmaf(P.r[1], hY, __builtin_fmaf(P.r[2], hZ, P.r[3])));
        const float hy = __builtin_fmaf(P.r[4], hX, __builtin_fmaf(P.r[5], hY, __builtin_fmaf(P.r[6], hZ, P.r[7])));
        flag = false;
        bool v;
        if constexpr (KIND == SEG_RING) {
            const float rr = __builtin_fmaf(hx, hx, hy * hy);
            v = (t > EPS_F) & (rr >= P.k0) & (rr <= P.k1);
        } else {
            v = (t > EPS_F) & (__builtin_fabsf(hx) <= 1.0f) & (__builtin_fabsf(hy) <= 1.0f);
        }
        return v ? t : finf();
    }
}

__device__ __forceinline__ float fast_test_any(const FastPrim &P, const Ray32 &R, bool &flag)
{
    switch (P.type) {
    case RF_CYLINDER: return fast_test<SEG_CYL>(P, R, flag);
    case RF_FRUSTUM: return fast_test<SEG_FRUSTUM>(P, R, flag);
    case RF_SPHERE: return fast_test<SEG_SPHERE>(P, R, flag);
    case RF_RECTANGLE: return fast_test<SEG_RECT>(P, R, flag);
    default: return fast_test<SEG_RING>(P, R, flag);
    }
}

__device__ __forceinline__ Ray64 widen(const Ray32 &r)
{
    Ray64 o;
    o.sx = r.sx; o.sy = r.sy; o.sz = r.sz; o.cx = r.cx; o.cy = r.cy; o.cz = r.cz;
    return o;
}

// FP64 re-solve of one flagged (ray, primitive) pair from the caller's
// record; the FP64 root is rounded back for the nearest-hit comparison.
// Kept out of line: it runs for a small fraction of pairs, and inlining it
// into the unrolled multi-ray loop would size the whole kernel's register
// allocation for FP64 temporaries of every ray at once.
__device__ __attribute__((noinline)) float refine_prim(CPTR(rf_primitive) P, int orig, const Ray32 &R)
{
    const rf_primitive p = ldc(P + orig);
    return (float)hit_d(p.type, p.invm, p.util1, p.util2, widen(R));
}

__device__ __attribute__((noinline)) float refine_bv(CPTR(rf_bounding_volume) B, int w, const Ray32 &R)
{
    const rf_bounding_volume b = ldc(B + w);
    return (float)hit_d(b.type, b.invm, b.util1, b.util2, widen(R));
}

// Deferred FP64 re-solves of the plain fast kernel: a flagged (ray, primitive)
// pair is kept out of the FP32 minimum and queued in a per-lane list in LDS
// (PEND_K entries per lane, strided by the workgroup size); after the scan the
// list is re-solved in FP64 and merged into the nearest hit.  A lane whose
// list overflows re-traces that ray entirely in FP64.  No FP64 code runs
// inside the primitive loop.
constexpr int PEND_K = 4;

struct Pending {
    uint32_t *buf;     // LDS, PEND_K * blockDim entries: (orig << 3) | ray slot
    int count;
    uint32_t overflow; // bit i: ray slot i needs a full FP64 re-trace
};

template <int KIND, int NR, bool REFINE>
__device__ __forceinline__ void seg_loop(CPTR(FastPrim) F, int begin, int end, const Ray32 (&R)[NR],
                                         float (&tb)[NR], int (&ib)[NR], Pending &pd)
{
    for (int w = begin; w < end; ++w) {
        const FastPrim q = ldc(F + w);
        float t[NR];
        bool f[NR];
        bool any = false;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            t[i] = fast_test<KIND>(q, R[i], f[i]);
            any |= f[i];
        }
        if constexpr (REFINE && KIND <= SEG_SPHERE) {
            if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
                for (int i = 0; i < NR; ++i) {
                    if (f[i]) {
                        if (pd.count < PEND_K) {
                            pd.buf[pd.count * blockDim.x + threadIdx.x] = ((uint32_t)q.orig << 3) | (uint32_t)i;
                            ++pd.count;
                        } else {
                            pd.overflow |= 1u << i;
                        }
                        t[i] = finf();
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < NR; ++i)
            if (t[i] < tb[i]) { tb[i] = t[i]; ib[i] = q.orig; }
    }
}

// FP64 stage of the deferred re-solves (see Pending).  The flagged ray is
// regenerated from its index (emission is a pure function of the random
// block), so the scan's ray registers are dead here.
template <int NR>
__device__ __forceinline__ void resolve_pending(const KParams &kp, CPTR(GeomDesc) G, const rf_primitive &E,
                                                uint32_t gid, int emitter, int64_t wi0, int64_t base,
                                                float (&tb)[NR], int (&ib)[NR], Pending &pd, uint32_t &nref)
{
    CPTR(rf_primitive) P = cst(G->prims);
    const int n = G->n;
    for (int e = 0; e < pd.count + __builtin_popcount(pd.overflow); ++e) {
        int prim_lo, prim_hi, k;
        if (e < pd.count) {
            const uint32_t ent = pd.buf[e * blockDim.x + threadIdx.x];
            prim_lo = (int)(ent >> 3);
            prim_hi = prim_lo + 1;
            k = (int)(ent & 7u);
        } else {
            // a slot whose list overflowed: re-trace the whole ray in FP64
            int m = e - pd.count;
            uint32_t ov = pd.overflow;
            while (m--) ov &= ov - 1;
            k = __builtin_ctz(ov);
            prim_lo = 0;
            prim_hi = n;
        }
        const int64_t rl = base + k;
        float f[4], r4[4];
        workitem_floats(kp, gid, emitter, wi0 + (rl >> 2), f);
        rotate_rands(f, (int)(rl & 3), r4);
        const Ray64 r64 = widen(emit_fast(&E, r4));
        float tk = finf();
        int ik = -1;
#pragma unroll
        for (int i = 0; i < NR; ++i)
            if (k == i) { tk = tb[i]; ik = ib[i]; }
        if (prim_hi - prim_lo > 1) { tk = finf(); ik = -1; }
        for (int w = prim_lo; w < prim_hi; ++w) {
            const rf_primitive p = ldc(P + w);
            const float t = (float)hit_d(p.type, p.invm, p.util1, p.util2, r64);
            if (t < tk || (t == tk && w < ik)) { tk = t; ik = w; }
        }
#pragma unroll
        for (int i = 0; i < NR; ++i)
            if (k == i) { tb[i] = tk; ib[i] = ik; }
        nref += (uint32_t)(prim_hi - prim_lo);
    }
}

template <int NR, bool REFINE>
__device__ __forceinline__ void scan_fast(CPTR(GeomDesc) G, const Ray32 (&R)[NR], float (&tb)[NR],
                                          int (&ib)[NR], Pending &pd)
{
    CPTR(FastPrim) F = cst(G->fast);
    seg_loop<SEG_CYL, NR, REFINE>(F, G->seg_begin[SEG_CYL], G->seg_end[SEG_CYL], R, tb, ib, pd);
    seg_loop<SEG_FRUSTUM, NR, REFINE>(F, G->seg_begin[SEG_FRUSTUM], G->seg_end[SEG_FRUSTUM], R, tb, ib, pd);
    seg_loop<SEG_SPHERE, NR, REFINE>(F, G->seg_begin[SEG_SPHERE], G->seg_end[SEG_SPHERE], R, tb, ib, pd);
    seg_loop<SEG_RING, NR, REFINE>(F, G->seg_begin[SEG_RING], G->seg_end[SEG_RING], R, tb, ib, pd);
    seg_loop<SEG_RECT, NR, REFINE>(F, G->seg_begin[SEG_RECT], G->seg_end[SEG_RECT], R, tb, ib, pd);
}

// Child scan of the fast BV kernel over the caller's order [start, end).
template <bool REFINE>
__device__ __forceinline__ int find_fast_range(CPTR(FastPrim) F, CPTR(rf_primitive) P, int start, int end,
                                               const Ray32 &R, uint32_t &nref)
{
    int obj = -1;
    float tbest = finf();
    for (int w = start; w < end; ++w) {
        const FastPrim q = ldc(F + w);
        bool f;
        float t = fast_test_any(q, R, f);
        if constexpr (REFINE) {
            if (__builtin_amdgcn_ballot_w64(f)) {
                if (f) { t = refine_prim(P, w, R); ++nref; }
            }
        }
        if (t < tbest) { tbest = t; obj = w; }
    }
    return obj;
}

template <bool REFINE>
__device__ __forceinline__ int find_bv_fast(CPTR(GeomDesc) G, const Ray32 &R, uint64_t &tests, uint32_t &nref)
{
    CPTR(FastPrim) F = cst(G->fast_orig);
    CPTR(FastPrim) FB = cst(G->fast_bvs);
    CPTR(BvMeta) M = cst(G->bv_meta);
    CPTR(rf_primitive) P = cst(G->prims);
    CPTR(rf_bounding_volume) B = cst(G->bvs);
    const int nbv = G->n_bvs;
    int bvIndex = -1, objectIndex = -1;
    float tbest = finf();
    for (int w = 0; w < nbv; ++w) {
        const FastPrim q = ldc(FB + w);
        const BvMeta meta = ldc(M + w);
        bool f;
        float t = fast_test_any(q, R, f);
        if constexpr (REFINE) {
            if (__builtin_amdgcn_ballot_w64(f)) {
                if (f) { t = refine_bv(B, w, R); ++nref; }
            }
        }
        ++tests;
        if (t < tbest) {
            if (meta.settings == 1) {
                objectIndex = find_fast_range<REFINE>(F, P, meta.start, meta.end, R, nref);
                tests += (uint64_t)(meta.end - meta.start);
                if (objectIndex != -1) { tbest = t; bvIndex = w; }
            } else {
                tbest = t; bvIndex = w;
            }
        }
    }
    bool pending = (bvIndex != -1 && objectIndex == -1);
    for (;;) {
        const uint64_t mask = __ballot(pending);
        if (mask == 0) break;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        const int bsel = __builtin_amdgcn_readlane(bvIndex, leader);
        if (pending && bvIndex == bsel) {
            const BvMeta meta = ldc(M + bsel);
            objectIndex = find_fast_range<REFINE>(F, P, meta.start, meta.end, R, nref);
            tests += (uint64_t)(meta.end - meta.start);
            pending = false;
        }
    }
    return objectIndex;
}

// ---------------------------------------------------------------------------
// Tally
// ---------------------------------------------------------------------------
template <bool TRACE, bool DIRECT>
__device__ __forceinline__ void tally(const KParams &kp, CPTR(GeomDesc) G, int emitter, int64_t ray_abs,
                                      int idx, uint32_t *bins)
{
    if constexpr (TRACE) {
        kp.trace_out[ray_abs] = idx;
    } else {
        if (idx < 0) return;
        if (DIRECT || kp.bins == 0)
            atomicAdd(&G->hits[(int64_t)emitter * G->n + idx], 1ull);
        else
            atomicAdd(&bins[idx], 1u);
    }
}

// ---------------------------------------------------------------------------
// Unit body: the rays of work items [lo, hi) (unit-local) handled by this
// thread, stepping by `stride` work items.
// ---------------------------------------------------------------------------
template <int MATH, bool BV, bool TRACE, int NR, bool REFINE, bool DIRECT>
__device__ __forceinline__ void run_unit(const KParams &kp, CPTR(GeomDesc) G, uint32_t gid, int emitter,
                                         int64_t wi0, int64_t n_wi, int64_t first, int64_t stride,
                                         uint32_t *bins, uint32_t *pend, uint64_t &tests, uint32_t &nref)
{
    const rf_primitive E = ldc(cst(G->prims) + emitter);
    const int64_t ray0 = (wi0 - kp.trace_wi0) * 4;   // absolute ray index base (trace launches)
    if constexpr (MATH == KM_FAST && !BV && !DIRECT) {
        // NR rays per thread: ray groups first, first+stride, ... of NR consecutive rays
        const int64_t n_rays = n_wi * 4;
        for (int64_t base = first * NR; base < n_rays; base += stride * NR) {
            Ray32 R[NR];
            float tb[NR];
            int ib[NR];
            float f[4];
            int64_t cur_wi = -1;
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                const int64_t rl = base + i;
                const int64_t wl = rl >> 2;
                if (wl != cur_wi) { workitem_floats(kp, gid, emitter, wi0 + wl, f); cur_wi = wl; }
                float r4[4];
                rotate_rands(f, (int)(rl & 3), r4);
                R[i] = emit_fast(&E, r4);
                tb[i] = finf();
                ib[i] = -1;
            }
            Pending pd{pend, 0, 0u};
            scan_fast<NR, REFINE>(G, R, tb, ib, pd);
            if constexpr (REFINE) {
                if (__builtin_amdgcn_ballot_w64(pd.count > 0 || pd.overflow != 0))
                    resolve_pending<NR>(kp, G, E, gid, emitter, wi0, base, tb, ib, pd, nref);
            }
#pragma unroll
            for (int i = 0; i < NR; ++i)
                if (base + i < n_rays) tally<TRACE, DIRECT>(kp, G, emitter, ray0 + base + i, ib[i], bins);
        }
    } else {
        // one work item (4 rays, one at a time) per thread and step
        for (int64_t wl = first; wl < n_wi; wl += stride) {
            float f[4];
            workitem_floats(kp, gid, emitter, wi0 + wl, f);
            for (int v = 0; v < 4; ++v) {
                float r4[4];
                rotate_rands(f, v, r4);
                int idx;
                if constexpr (MATH == KM_COMPAT) {
                    const Ray32 ray = emit_c(&E, r4);
                    idx = BV ? find_bv_cd<KM_COMPAT>(cst(G->prims), cst(G->bvs), G->n_bvs, ray, tests)
                             : find_c(cst(G->prims), 0, G->n, ray);
                } else if constexpr (MATH == KM_FP64) {
                    const double d4[4] = {r4[0], r4[1], r4[2], r4[3]};
                    const Ray64 ray = emit_d(&E, d4);
                    idx = BV ? find_bv_cd<KM_FP64>(cst(G->prims), cst(G->bvs), G->n_bvs, ray, tests)
                             : find_d(cst(G->prims), 0, G->n, ray);
                } else {
                    const Ray32 ray = emit_fast(&E, r4);
                    if constexpr (BV) {
                        idx = find_bv_fast<REFINE>(G, ray, tests, nref);
                    } else {
                        Ray32 R1[1] = {ray};
                        float tb1[1] = {finf()};
                        int ib1[1] = {-1};
                        Pending pd{pend, 0, 0u};
                        scan_fast<1, REFINE>(G, R1, tb1, ib1, pd);
                        if constexpr (REFINE) {
                            if (__builtin_amdgcn_ballot_w64(pd.count > 0 || pd.overflow != 0))
                                resolve_pending<1>(kp, G, E, gid, emitter, wi0, wl * 4 + v, tb1, ib1, pd, nref);
                        }
                        idx = ib1[0];
                    }
                }
                tally<TRACE, DIRECT>(kp, G, emitter, ray0 + wl * 4 + v, idx, bins);
            }
        }
    }
}

__device__ __forceinline__ void flush_stats(const KParams &kp, uint64_t tests, uint32_t nref)
{
    __shared__ unsigned long long s_stats[2];
    if (threadIdx.x == 0) { s_stats[0] = 0; s_stats[1] = 0; }
    __syncthreads();
    if (tests) atomicAdd(&s_stats[0], (unsigned long long)tests);
    if (nref) atomicAdd(&s_stats[1], (unsigned long long)nref);
    __syncthreads();
    if (threadIdx.x == 0 && kp.stats) {
        if (s_stats[0]) atomicAdd(&kp.stats[0], s_stats[0]);
        if (s_stats[1]) atomicAdd(&kp.stats[1], s_stats[1]);
    }
}

// Persistent queue kernel: every workgroup pulls units until the queue is
// empty (all waves reach the same exit).
template <int MATH, bool BV, bool TRACE, int NR, bool REFINE>
__global__ void __launch_bounds__(256) k_queue(KParams kp)
{
    // dynamic LDS: [bins: kp.bins x u32][pending re-solves: PEND_K x blockDim x u32]
    extern __shared__ uint32_t smem[];
    uint32_t *bins = smem;
    uint32_t *pend = smem + (TRACE ? 0 : kp.bins);
    __shared__ int s_unit;
    uint64_t tests = 0;
    uint32_t nref = 0;
    const bool lds_tally = !TRACE && kp.bins > 0;
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int)atomicAdd(kp.queue_head, 1u);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= kp.n_units) break;
        CPTR(GeomDesc) G;
        int emitter;
        int64_t wi0, n_wi;
        uint32_t gid;
        if (kp.units) {
            const WorkUnit U = ldc(cst(kp.units) + u);
            G = cst(kp.geoms) + U.geom;
            emitter = U.emitter; wi0 = U.wi0; n_wi = U.n_wi;
            gid = G->geometry_id;
        } else {
            G = cst(kp.geoms);
            emitter = kp.imp_emitter;
            wi0 = kp.imp_wi0 + (int64_t)u * kp.imp_unit;
            const int64_t left = kp.imp_wi0 + kp.imp_n_wi - wi0;
            n_wi = left < kp.imp_unit ? left : kp.imp_unit;
            gid = kp.imp_geom_id;
        }
        const int n = G->n;
        if (lds_tally) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) bins[i] = 0u;
            __syncthreads();
        }
        run_unit<MATH, BV, TRACE, NR, REFINE, false>(kp, G, gid, emitter, wi0, n_wi, threadIdx.x,
                                                     blockDim.x, bins, pend, tests, nref);
        if (lds_tally) {
            __syncthreads();
            unsigned long long *row = G->hits + (int64_t)emitter * n;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t c = bins[i];
                if (c) atomicAdd(&row[i], (unsigned long long)c);
            }
        }
        __syncthreads();
    }
    flush_stats(kp, tests, nref);
}

// Reference dispatch shape (RayFactorGPU.cl:481 launched as main.cpp:172-176):
// one work item per thread over the implicit unit, one global atomic per
// hitting ray.
template <int MATH, bool BV>
__global__ void __launch_bounds__(256) k_direct(KParams kp)
{
    extern __shared__ uint32_t smem[];   // pending re-solves: PEND_K x blockDim x u32
    CPTR(GeomDesc) G = cst(kp.geoms);
    uint64_t tests = 0;
    uint32_t nref = 0;
    const int64_t wl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    run_unit<MATH, BV, false, 1, true, true>(kp, G, kp.imp_geom_id, kp.imp_emitter, kp.imp_wi0, kp.imp_n_wi,
                                             wl, kp.imp_n_wi + 1, nullptr, smem, tests, nref);
    flush_stats(kp, tests, nref);
}

__global__ void k_fill_u32(unsigned int *p, unsigned int v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

hipError_t launch_fill_u32(unsigned int *p, unsigned int value, size_t count, hipStream_t stream)
{
    if (!count) return hipSuccess;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, stream, p, value, count);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, size_t count, hipStream_t stream)
{
    if (!count) return hipSuccess;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, stream, p, count);
    return hipGetLastError();
}

template <int MATH, bool BV, bool TRACE, int NR, bool REFINE>
static hipError_t go_queue(const LaunchCfg &c, const KParams &kp, hipStream_t s)
{
    hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
    return hipGetLastError();
}

template <int MATH, bool BV, bool TRACE>
static hipError_t go_queue_math(const LaunchCfg &c, const KParams &kp, hipStream_t s)
{
    if constexpr (MATH != KM_FAST) {
        return go_queue<MATH, BV, TRACE, 1, false>(c, kp, s);
    } else if constexpr (BV) {
        return c.refine ? go_queue<MATH, true, TRACE, 1, true>(c, kp, s)
                        : go_queue<MATH, true, TRACE, 1, false>(c, kp, s);
    } else {
        const bool r = c.refine != 0;
        switch (c.nr) {
        case 1: return r ? go_queue<MATH, false, TRACE, 1, true>(c, kp, s) : go_queue<MATH, false, TRACE, 1, false>(c, kp, s);
        case 2: return r ? go_queue<MATH, false, TRACE, 2, true>(c, kp, s) : go_queue<MATH, false, TRACE, 2, false>(c, kp, s);
        case 8: return r ? go_queue<MATH, false, TRACE, 8, true>(c, kp, s) : go_queue<MATH, false, TRACE, 8, false>(c, kp, s);
        default: return r ? go_queue<MATH, false, TRACE, 4, true>(c, kp, s) : go_queue<MATH, false, TRACE, 4, false>(c, kp, s);
        }
    }
}

template <int MATH>
static hipError_t go_math(const LaunchCfg &c, const KParams &kp, hipStream_t s)
{
    if (c.direct) {
        if (c.use_bv) hipLaunchKernelGGL((k_direct<MATH, true>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
        else hipLaunchKernelGGL((k_direct<MATH, false>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
        return hipGetLastError();
    }
    if (c.use_bv) return c.trace ? go_queue_math<MATH, true, true>(c, kp, s) : go_queue_math<MATH, true, false>(c, kp, s);
    return c.trace ? go_queue_math<MATH, false, true>(c, kp, s) : go_queue_math<MATH, false, false>(c, kp, s);
}

hipError_t launch_solve(const LaunchCfg &c, const KParams &kp, hipStream_t s)
{
    switch (c.math) {
    case KM_COMPAT: return go_math<KM_COMPAT>(c, kp, s);
    case KM_FP64: return go_math<KM_FP64>(c, kp, s);
    default: return go_math<KM_FAST>(c, kp, s);
    }
}

}  // namespace rf
