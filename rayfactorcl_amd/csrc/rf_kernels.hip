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
// never diverges.  The hot loops copy each record to VGPRs once (an SGPR
// operand would keep the arithmetic from dual-issuing).
//
//   compat  FP32 reference semantics, caller's primitive order, the 4 rays of
//           a work item per primitive visit
//   fp64    FP64 physics semantics, caller's primitive order
//   fast    FP32 with hardware rcp/sqrt over type-segmented primitive
//           records, NR rays per thread, FP64 re-solve of flagged roots;
//           variants with the ray transform on the matrix cores
//           (k_queue_mfma) and with rays binned by bounding surface
//           (k_queue_bvbin)
//
// "direct" launches keep the reference's dispatch shape instead: one work
// item per thread, 128-wide groups, one global atomic per hitting ray.
#include <algorithm>

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

// First position w of [b, e) (e - b >= 2) at which pass(w, i) holds, for each
// ray slot i, where pass is false on a prefix of the positions and true on the
// rest (e if it never holds): binary lifting, a wave-uniform number of steps
// with one probe per ray each.
template <int NR, class Pass>
__device__ __forceinline__ void band_lower_bound(int b, int e, Pass pass, int (&pos)[NR])
{
    int last[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) last[i] = b - 1;
    for (int step = 1 << (31 - __builtin_clz(e - b)); step > 0; step >>= 1) {
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int q = last[i] + step;
            const bool inside = q < e;
            last[i] = (inside && !pass(inside ? q : e - 1, i)) ? q : last[i];
        }
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) pos[i] = last[i] + 1;
}

// band_lower_bound from a first guess g[i] in [b, e] per ray slot: the guess
// is the answer when pass holds at g and fails just before it (positions b-1
// and e count as failing and holding); both probes are issued together.
// Slots whose guess is off fall back to the binary search.
template <int NR, class Pass>
__device__ __forceinline__ void band_lower_bound_guess(int b, int e, const int (&g)[NR], Pass pass,
                                                       int (&pos)[NR])
{
    bool off = false;
    bool ok[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int gi = g[i];
        const bool at = pass(min(gi, e - 1), i);
        const bool before = pass(max(gi - 1, b), i);
        ok[i] = (gi >= e || at) && (gi <= b || !before);
        pos[i] = gi;
        off |= !ok[i];
    }
    if (__builtin_amdgcn_ballot_w64(off)) {
        int bs[NR];
        band_lower_bound<NR>(b, e, pass, bs);
#pragma unroll
        for (int i = 0; i < NR; ++i) pos[i] = ok[i] ? pos[i] : bs[i];
    }
}

// First guess of a lower bound over an evenly spaced stack: the position
// whose offset first reaches `target` (ASC: offsets increasing, pass where
// offset >= target; else decreasing, pass where offset <= target), from the
// first offset r0 and inv = bands per unit offset; a NaN target guesses e.
__device__ __forceinline__ int band_guess(int b, int e, float r0, float inv, float target)
{
    float f = __builtin_ceilf((target - r0) * inv);
    f = __builtin_fminf(__builtin_fmaxf(f, 0.0f), (float)(e - b));
    return target != target ? e : b + (int)f;
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

// FP64 band stack [b, e) in the caller's order (the compat stacks: consecutive
// cylinders whose invm differs only in invm[11]).  hit_d's cylinder
// (quad_select_d) has the same A, B, C, roots and o.cz * t for every band, and
// o.sz = ((c0 sx + c1 sy) + c2 sz) + c3 differs by the last term only.  Per
// band it returns the lower root when it lies past EPS with z in [0, 1], else
// the higher root under the same test, else inf; with lo <= hi the stack's
// nearest hit is lo at the first band holding z(lo), else hi at the first band
// holding z(hi) (bands walked last to first, as stack_loop).  A root that is
// absent or not past EPS becomes a NaN product, which fails every range test.
__device__ __forceinline__ uint32_t stack_d(CPTR(rf_primitive) P, CPTR(BandRec) BR, int b, int e, int mode,
                                            float inv, const Ray64 &w, double &tbest, int &obj, uint64_t &ev)
{
    const rf_primitive hs = ldc(P + b);
    const float *iv = hs.invm;
    const double a0 = iv[0], a1 = iv[1], a2 = iv[2], a3 = iv[3];
    const double b0 = iv[4], b1 = iv[5], b2 = iv[6], b3 = iv[7];
    const double c0 = iv[8], c1 = iv[9], c2 = iv[10];
    const double sx = ((a0 * w.sx + a1 * w.sy) + a2 * w.sz) + a3;
    const double cx = (a0 * w.cx + a1 * w.cy) + a2 * w.cz;
    const double sy = ((b0 * w.sx + b1 * w.sy) + b2 * w.sz) + b3;
    const double cy = (b0 * w.cx + b1 * w.cy) + b2 * w.cz;
    const double S = (c0 * w.sx + c1 * w.sy) + c2 * w.sz;
    const double cz = (c0 * w.cx + c1 * w.cy) + c2 * w.cz;
    const double A = cx * cx + cy * cy;
    const double B = sx * cx + sy * cy;
    const double C = (sx * sx + sy * sy) - 1.0;
    const double qnan = __builtin_nan("");
    double r1 = qnan, r2 = qnan;
    if (A == 0.0) {
        if (B != 0.0) r1 = r2 = -C / (2.0 * B);
    } else {
        const double D = B * B - A * C;
        if (!(D < 0.0)) {
            const double sq = __builtin_sqrt(D);
            const double q = -(B + (B >= 0.0 ? sq : -sq));
            if (q != 0.0) {
                r1 = q / A;
                r2 = C / q;
            }
        }
    }
    const double lo = r1 < r2 ? r1 : r2;
    const double hi = r1 < r2 ? r2 : r1;
    const double plo = lo > EPS_D ? cz * lo : qnan;
    const double phi = hi > EPS_D ? cz * hi : qnan;
    if (mode >= 2) {
        // sorted band offsets: z at each root is monotone over the positions
        // (an add is non-decreasing in each operand), so the first band
        // holding it is the first position past the range's lower (mode 2)
        // or upper (mode 3) end, if z there is in range
        const bool asc = mode == 2;
        uint32_t probes = 0;
        const auto zat = [&](int k, double p) { ++probes; return (S + (double)BR[k].r11) + p; };
        const float r0 = ldc(BR + b).r11;
        const double lim = asc ? 0.0 : 1.0;
        const auto pass = [&](double z) { return asc ? z >= 0.0 : z <= 1.0; };
        // the first position past the range's end for product p, and whether
        // z there is in range; the guess and the position before it are
        // probed together
        const auto root_search = [&](double p, int &pos) {
            const int g = band_guess(b, e, r0, inv, (float)(lim - (S + p)));
            const double zg = zat(min(g, e - 1), p);
            const double zm = zat(max(g - 1, b), p);
            double z = zg;
            pos = g;
            if (!((g >= e || pass(zg)) && (g <= b || !pass(zm)))) {
                int bs[1];
                band_lower_bound<1>(b, e, [&](int k, int) { return pass(zat(k, p)); }, bs);
                pos = bs[0];
                z = zat(min(pos, e - 1), p);
            }
            return pos < e && z >= 0.0 && z <= 1.0;
        };
        int pl, ph;
        const bool vlo = root_search(plo, pl);
        const bool vhi = root_search(phi, ph);
        const double c = vlo ? lo : (vhi ? hi : dinf());
        if (c < tbest) { tbest = c; obj = BR[min(vlo ? pl : ph, e - 1)].orig; }
        ev += 1u + probes;
        return kFlopsStackD + probes * (kFlopsBandD - 1u);
    }
    int ilo = -1, ihi = -1;
    CPTR(BandRec) q = BR + (e - 1);
    BandRec next = ldc(q);
    for (int wb = e - 1; wb >= b; --wb) {
        const double c3 = next.r11;
        const int orig = next.orig;
        --q;
        next = ldc(q);
        asm volatile("" ::: "memory");
        const double sz = S + c3;
        const double zlo = sz + plo;
        const double zhi = sz + phi;
        ilo = (zlo >= 0.0 && zlo <= 1.0) ? orig : ilo;
        ihi = (zhi >= 0.0 && zhi <= 1.0) ? orig : ihi;
    }
    const double c = ilo >= 0 ? lo : (ihi >= 0 ? hi : dinf());
    const int ci = ilo >= 0 ? ilo : ihi;
    if (c < tbest) { tbest = c; obj = ci; }
    ev += (uint64_t)(1 + e - b);
    return kFlopsStackD + (uint32_t)(e - b) * kFlopsBandD;
}

// findObject in the FP64 form over the caller's range [s, e) with the compat
// runs when `stacks`: a stack cut by the range keeps the bands inside it (two
// or more stay a stack, and a sorted stack stays sorted); adds the stacks'
// algorithmic flops to `sflops` and the evaluations of the stacks -- of the
// whole range when `count_all` -- to `ev`.
__device__ __forceinline__ int find_d_runs(CPTR(GeomDesc) G, int s, int e, const Ray64 &ray, int stacks,
                                           uint64_t &sflops, uint64_t &ev, bool count_all = false)
{
    CPTR(rf_primitive) P = cst(G->prims);
    if (!stacks || s >= e) {
        if (count_all) ev += (uint64_t)(e > s ? e - s : 0);
        return find_d(P, s, e, ray);
    }
    int obj = -1;
    double tbest = dinf();
    const int n_runs = G->n_c_runs;
    for (int j = cst(G->c_run_of)[s]; j < n_runs; ++j) {
        const CylRun run = ldc(cst(G->c_runs) + j);
        if (run.begin >= e) break;
        const int rb = max(run.begin, s), re = min(run.end, e);
        if (run.stack && re - rb >= 2) {
            sflops += stack_d(P, cst(G->c_bands), rb, re, stacks >= 2 ? run.stack : 1, run.inv, ray, tbest, obj, ev);
        } else {
            if (count_all) ev += (uint64_t)(re - rb);
            for (int w = rb; w < re; ++w) {
                const rf_primitive p = ldc(P + w);
                const double t = hit_d(p.type, p.invm, p.util1, p.util2, ray);
                if (t < tbest) { tbest = t; obj = w; }
            }
        }
    }
    return obj;
}

// ---------------------------------------------------------------------------
// fast: per-type tests over FastPrim records.  Each returns t (+INF on a
// miss) and sets flag when the FP32 roots are ill-conditioned.
// ---------------------------------------------------------------------------
// A VALU instruction with an SGPR data operand does not dual-issue on gfx950
// (measured: 15 vs 28 lane-FMAs per clock per SIMD).  Wave-uniform operands of
// the hot loops (primitive rows, thresholds) are therefore copied into VGPRs
// once per primitive / per kernel and the per-ray arithmetic reads them there.
__device__ __forceinline__ float vcopy(float x)
{
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

// Two floats per v_mov_b64 (gfx950 has the 64-bit move).
__device__ __forceinline__ void vcopy2(float a, float b, float &ra, float &rb)
{
    const uint64_t x = ((uint64_t)__float_as_uint(b) << 32) | __float_as_uint(a);
    uint64_t r;
    asm("v_mov_b64 %0, %1" : "=v"(r) : "s"(x));
    ra = __uint_as_float((uint32_t)r);
    rb = __uint_as_float((uint32_t)(r >> 32));
}

__device__ __forceinline__ FastPrim vcopy(const FastPrim &q)
{
    FastPrim v;
#pragma unroll
    for (int i = 0; i < 12; i += 2) vcopy2(q.r[i], q.r[i + 1], v.r[i], v.r[i + 1]);
    vcopy2(q.k0, q.k1, v.k0, v.k1);
    v.orig = q.orig;
    v.type = q.type;
    return v;
}

// One primitive type against the NR rays of find_c_multi: FastOps first; a
// ray whose operands left FastOps' exact range repeats the test with IeeeOps.
template <int TYPE, int NR>
__device__ __forceinline__ void scan_c_type(const float (&m)[12], float u1, float u2, const Ray32 (&ray)[NR],
                                            float (&t)[NR])
{
    bool redo[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        FastOps ops;
        t[r] = hit_c_ops<TYPE>(m, u1, u2, ray[r], ops);
        redo[r] = !ops.ok;
        any |= redo[r];
    }
    if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (redo[r]) t[r] = hit_c_t<TYPE>(m, u1, u2, ray[r]);
    }
}

// Compat scan of the NR rays of one work item over records [start, end):
// each record is read once, copied to VGPRs and tested against every ray with
// the reference's arithmetic (hit_c_ops, IEEE-exact); caller order and strict
// < keep the reference's tie rule (RayFactorGPU.cl:445-472), so each ray's
// result is find_c's.  tbest / obj carry the running nearest hit.
template <int NR>
__device__ __forceinline__ void scan_c_range(CPTR(rf_primitive) P, int start, int end, const Ray32 (&ray)[NR],
                                             float (&tbest)[NR], int (&obj)[NR])
{
    for (int w = start; w < end; ++w) {
        const rf_primitive p = ldc(P + w);
        float m[12];
#pragma unroll
        for (int k = 0; k < 12; k += 2) vcopy2(p.invm[k], p.invm[k + 1], m[k], m[k + 1]);
        float t[NR];
        switch (p.type) {
        case RF_CYLINDER: scan_c_type<RF_CYLINDER, NR>(m, 0.f, 0.f, ray, t); break;
        case RF_FRUSTUM: scan_c_type<RF_FRUSTUM, NR>(m, vcopy(p.util1), 0.f, ray, t); break;
        case RF_SPHERE: scan_c_type<RF_SPHERE, NR>(m, 0.f, 0.f, ray, t); break;
        case RF_ANNULUS: scan_c_type<RF_ANNULUS, NR>(m, vcopy(p.util1), vcopy(p.util2), ray, t); break;
        case RF_RECTANGLE: scan_c_type<RF_RECTANGLE, NR>(m, 0.f, 0.f, ray, t); break;
        case RF_DISC: scan_c_type<RF_DISC, NR>(m, 0.f, 0.f, ray, t); break;
        default:
#pragma unroll
            for (int r = 0; r < NR; ++r) t[r] = finf();
            break;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (t[r] < tbest[r]) { tbest[r] = t[r]; obj[r] = w; }
    }
}

// Band loop of stack_c, bands e-1 down to b (AXIAL: wave-uniform choice of the
// one-FMA z origin).
template <int NR, bool AXIAL>
__device__ __forceinline__ void stack_c_bands(const float (&m)[12], CPTR(BandRec) BR, int b, int e,
                                              const Ray32 (&ray)[NR], const float (&P1)[NR], const float (&P2)[NR],
                                              int (&I1)[NR], int (&I2)[NR])
{
    // BR[-1] exists (pad record): the next record's load needs no clamp
    CPTR(BandRec) q = BR + (e - 1);
    BandRec next = ldc(q);
    for (int wb = e - 1; wb >= b; --wb) {
        float m11, of;
        vcopy2(next.r11, __int_as_float(next.orig), m11, of);
        const int orig = __float_as_int(of);
        --q;
        next = ldc(q);
        // keeps the next record's load here, a band ahead of its use
        asm volatile("" ::: "memory");
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const Ray32 &w = ray[r];
            const float sz = AXIAL ? __builtin_fmaf(m[10], w.sz, m11)
                                   : __builtin_fmaf(m[8], w.sx, __builtin_fmaf(m[9], w.sy, __builtin_fmaf(m[10], w.sz, m11)));
            const float z1 = sz + P1[r];
            const float z2 = sz + P2[r];
            const bool v1 = __builtin_amdgcn_fmed3f(z1, 0.0f, 1.0f) == z1;     // z1 in [0, 1]
            const bool v2 = (z2 > 0.0f) & (z2 < 1.0f);
            I1[r] = v1 ? orig : I1[r];
            I2[r] = v2 ? orig : I2[r];
        }
    }
}

// stack_c over a stack whose band offsets invm[11] are sorted over the
// positions (asc: non-decreasing, else non-increasing; all finite).  z at
// each root, o.sz(w) + o.cz t with o.sz(w) an FMA chain ending in invm[11](w),
// is then monotone over the positions (FMA and add are non-decreasing in
// their addends), so the bands holding root 1 (z in [0, 1]) or root 2
// (z in (0, 1)) form one run each, and its first position is the first past
// the range's lower (asc) or upper end -- if z there is in range.  I1 / I2
// return positions (-1: none); `probes` counts the z evaluations.
template <int NR, bool AXIAL>
__device__ __forceinline__ void stack_c_search(const float (&m)[12], CPTR(BandRec) BR, int b, int e, bool asc,
                                               float inv, const Ray32 (&ray)[NR], const float (&P1)[NR], const float (&P2)[NR],
                                               int (&I1)[NR], int (&I2)[NR], uint32_t &probes)
{
    const auto zat = [&](int w, int r, float p) {
        const Ray32 &q = ray[r];
        const float m11 = BR[w].r11;
        const float sz = AXIAL ? __builtin_fmaf(m[10], q.sz, m11)
                               : __builtin_fmaf(m[8], q.sx, __builtin_fmaf(m[9], q.sy, __builtin_fmaf(m[10], q.sz, m11)));
        ++probes;
        return sz + p;
    };
    // first guesses: z is about o.sz with invm[11] = 0, plus the root's
    // product, plus invm[11]
    const float r0 = ldc(BR + b).r11;
    float s0[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const Ray32 &q = ray[r];
        s0[r] = AXIAL ? m[10] * q.sz : __builtin_fmaf(m[8], q.sx, __builtin_fmaf(m[9], q.sy, m[10] * q.sz));
    }
    // one root: the first position where z passes the range's lower (asc) or
    // upper end, and z there; the guess and the position before it are
    // probed together, and z at the guess is the answer's when it holds
    const auto root_search = [&](const float (&P)[NR], bool open, int (&pos)[NR], float (&zp)[NR]) {
        const auto pass = [&](float z) {
            return asc ? (open ? z > 0.0f : z >= 0.0f) : (open ? z < 1.0f : z <= 1.0f);
        };
        bool ok[NR];
        bool off = false;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int g = band_guess(b, e, r0, inv, (asc ? 0.0f : 1.0f) - (s0[r] + P[r]));
            const float zg = zat(min(g, e - 1), r, P[r]);
            const float zm = zat(max(g - 1, b), r, P[r]);
            ok[r] = (g >= e || pass(zg)) && (g <= b || !pass(zm));
            pos[r] = g;
            zp[r] = zg;
            off |= !ok[r];
        }
        if (__builtin_amdgcn_ballot_w64(off)) {
            int bs[NR];
            band_lower_bound<NR>(b, e, [&](int w, int r) { return pass(zat(w, r, P[r])); }, bs);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (!ok[r]) {
                    pos[r] = bs[r];
                    zp[r] = zat(min(bs[r], e - 1), r, P[r]);
                }
            }
        }
    };
    int p1[NR], p2[NR];
    float z1[NR], z2[NR];
    root_search(P1, false, p1, z1);
    root_search(P2, true, p2, z2);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const bool v1 = p1[r] < e && __builtin_amdgcn_fmed3f(z1[r], 0.0f, 1.0f) == z1[r];
        const bool v2 = p2[r] < e && (z2[r] > 0.0f) & (z2[r] < 1.0f);
        I1[r] = v1 ? p1[r] : -1;
        I2[r] = v2 ? p2[r] : -1;
    }
}

// Compat band stack [b, e) in the caller's order: consecutive cylinders whose
// invm rows are bitwise equal except invm[11].  cylinderHit (:189-222, here
// hit_c_ops / roots_select_c) then computes the same A, B, C, discriminant,
// roots and o.cz * t for every band; a band differs only in o.sz.  Per band
// the reference returns t1 when t1 >= EPS and z1 in [0, 1], else t2 when
// t2 < inf, t2 > EPS and z2 in (0, 1) (its t2 < t1 test can only pass once t1
// is set to inf: 1/A >= 0 orders t1 <= t2), else inf; a NaN t1 is returned as
// it is and never wins.  With t1 <= t2 (and z1 == z2 when they are equal) the
// stack's nearest hit is t1 at the first band holding z1, else t2 at the first
// band holding z2, found walking the bands last to first as stack_loop does.
// The root conditions are folded into NaN products, which fail every range
// test.  An axial z row (invm[8] == invm[9] == 0) takes one FMA for o.sz: the
// two dropped FMAs add a signed zero, which no range test can tell apart.
// mode 1 walks every band; 2 and 3 (band offsets sorted, non-decreasing or
// non-increasing) search them (stack_c_search).  [b, e) is the part of the
// run scanned (all of it, or a bounding surface's children).
// Analytic positions (run.an_margin > 0, axial): as in stack_loop, each root
// is first placed in the run's progression of offsets in FP64.  Band w's z
// is (m[10] s_z + invm[11](w)) + o.cz t, two roundings, so with Z* its exact
// value and p = o.cz t, |z - Z*| <= E = 2^-24 (|p| + 4.01) where |Z*| <= 2;
// a root whose distance dc from the nearest position of [b, e) satisfies
// |delta| dc <= 1/2 - m (m = dev + 2E + rounding) lies on that band with
// z in (0, 1) and on no other, and one with |delta| dc >= 1/2 + m on none:
// both range tests (z1 in [0, 1], z2 in (0, 1)) are decided, and the rays so
// placed skip the search (their products become NaN there).
// Returns the stack's algorithmic flops summed over the NR rays.
// ev gains the stack's evaluations over the NR rays: a quadric per ray and a
// z-range check per band looked at (one per placed root).
template <int NR>
__device__ __forceinline__ uint32_t stack_c(CPTR(rf_primitive) P, CPTR(BandRec) BR, const CylRun &run, int b, int e,
                                            int mode, const Ray32 (&ray)[NR], float (&tbest)[NR], int (&obj)[NR],
                                            uint32_t &ev)
{
    const float inv = run.inv;
    const rf_primitive hs = ldc(P + b);
    float m[12];
#pragma unroll
    for (int k = 0; k < 12; k += 2) vcopy2(hs.invm[k], hs.invm[k + 1], m[k], m[k + 1]);
    const bool axial = ((__float_as_uint(hs.invm[8]) | __float_as_uint(hs.invm[9])) << 1) == 0u;
    const float qnan = __builtin_nanf("");
    float T1[NR], T2[NR], P1[NR], P2[NR];
    int I1[NR], I2[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const Ray32 &w = ray[r];
        const float sx = __builtin_fmaf(m[0], w.sx, __builtin_fmaf(m[1], w.sy, __builtin_fmaf(m[2], w.sz, m[3])));
        const float cx = __builtin_fmaf(m[0], w.cx, __builtin_fmaf(m[1], w.cy, m[2] * w.cz));
        const float sy = __builtin_fmaf(m[4], w.sx, __builtin_fmaf(m[5], w.sy, __builtin_fmaf(m[6], w.sz, m[7])));
        const float cy = __builtin_fmaf(m[4], w.cx, __builtin_fmaf(m[5], w.cy, m[6] * w.cz));
        const float cz = __builtin_fmaf(m[8], w.cx, __builtin_fmaf(m[9], w.cy, m[10] * w.cz));
        const float A = __builtin_fmaf(cx, cx, cy * cy);
        const float B = -__builtin_fmaf(sx, cx, sy * cy);
        const float C = __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, -1.f));
        const float discrim = sqrt_ieee(B * B - A * C);
        const float rA = recip_ieee(A);
        const float t1 = rA * (B - discrim);
        const float t2 = rA * (B + discrim);
        float p1 = cz * t1, p2 = cz * t2;
        const bool nan1 = t1 != t1;                              // no band returns a usable t
        p1 = (nan1 || t1 < EPS_F) ? qnan : p1;
        p2 = (nan1 || !(t2 < finf()) || !(t2 > EPS_F)) ? qnan : p2;
        T1[r] = t1;
        T2[r] = t2;
        P1[r] = p1;
        P2[r] = p2;
        asm volatile("" : "+v"(P1[r]), "+v"(P2[r]));
        I1[r] = -1;
        I2[r] = -1;
    }
    if (mode >= 2) {
        uint32_t probes = 0, placed = 0;
        bool sure[NR];
        int K1[NR], K2[NR];
        bool rest = true;
        if (axial && run.an_margin > 0.0f) {
            const float mfix = run.an_margin + 0x1.1p-21f;   // 2^-23 x 4.25
            const float ha = 0.5f * run.an_ainv;
            const double klo = (double)(b - run.begin), khi = (double)(e - 1 - run.begin);
            bool left = false;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                // z of the root on the run's first band, less the range's centre
                const double zb = __builtin_fma((double)m[10], (double)ray[r].sz, (double)run.r0) - 0.5;
                const auto place = [&](float p, bool &sr) {
                    const double x = (zb + (double)p) * run.an_ninv;
                    const double k = __builtin_fmin(__builtin_fmax(__builtin_rint(x), klo), khi);
                    const float dc = (float)__builtin_fabs(x - k);
                    const float mm = __builtin_fmaf(0x1.002p-23f, __builtin_fabsf(p), mfix);
                    const bool on = dc <= __builtin_fmaf(-mm, run.an_ainv, ha);
                    sr = on || !(dc < __builtin_fmaf(mm, run.an_ainv, ha));
                    return on ? run.begin + (int)k : -1;
                };
                bool s1, s2;
                K1[r] = place(P1[r], s1);
                K2[r] = place(P2[r], s2);
                sure[r] = s1 && s2;
                if (sure[r]) {
                    P1[r] = qnan;
                    P2[r] = qnan;
                }
                left |= !sure[r];
            }
            placed = 2 * NR;
            rest = __builtin_amdgcn_ballot_w64(left);
        } else {
#pragma unroll
            for (int r = 0; r < NR; ++r) sure[r] = false;
        }
        if (rest) {
            if (axial) stack_c_search<NR, true>(m, BR, b, e, mode == 2, inv, ray, P1, P2, I1, I2, probes);
            else stack_c_search<NR, false>(m, BR, b, e, mode == 2, inv, ray, P1, P2, I1, I2, probes);
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (sure[r]) {
                I1[r] = K1[r];
                I2[r] = K2[r];
            }
            const bool h1 = I1[r] >= 0;
            const float c = h1 ? T1[r] : (I2[r] >= 0 ? T2[r] : finf());
            if (c < tbest[r]) { tbest[r] = c; obj[r] = BR[min(h1 ? I1[r] : I2[r], e - 1)].orig; }
        }
        ev += NR + probes + placed;
        return NR * kFlopsStackC + (probes + placed) * ((axial ? kFlopsBandZ0Axial : kFlopsBandZ0) + 1u);
    }
    if (axial) stack_c_bands<NR, true>(m, BR, b, e, ray, P1, P2, I1, I2);
    else stack_c_bands<NR, false>(m, BR, b, e, ray, P1, P2, I1, I2);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const bool h1 = I1[r] >= 0;
        const float c = h1 ? T1[r] : (I2[r] >= 0 ? T2[r] : finf());
        const int ci = h1 ? I1[r] : I2[r];
        if (c < tbest[r]) { tbest[r] = c; obj[r] = ci; }
    }
    ev += NR * (uint32_t)(1 + e - b);
    return NR * (kFlopsStackC + (uint32_t)(e - b) * ((axial ? kFlopsBandZ0Axial : kFlopsBandZ0) + 2u));
}

// The compat scan of one work item's NR rays over the whole scene: band stacks
// (G->c_runs) when `stacks` (2: sorted stacks searched), else one per-test
// range.  Returns the stacks' algorithmic flops summed over the NR rays; ev
// gains their evaluations.
template <int NR>
__device__ __forceinline__ uint32_t find_c_multi(CPTR(GeomDesc) G, const Ray32 (&ray)[NR], int (&obj)[NR], int stacks,
                                                 uint32_t &ev)
{
    CPTR(rf_primitive) P = cst(G->prims);
    float tbest[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) { tbest[r] = finf(); obj[r] = -1; }
    uint32_t sf = 0;
    const int n_runs = stacks ? G->n_c_runs : 1;
    for (int j = 0; j < n_runs; ++j) {
        CylRun run{0, G->n, 0, 0};
        if (stacks) run = ldc(cst(G->c_runs) + j);
        if (run.stack)
            sf += stack_c<NR>(P, cst(G->c_bands), run, run.begin, run.end, stacks >= 2 ? run.stack : 1, ray, tbest,
                              obj, ev);
        else
            scan_c_range<NR>(P, run.begin, run.end, ray, tbest, obj);
    }
    return sf;
}

// The compat child scan of one ray over the caller's range [s, e): the
// compat runs cut to the range when `stacks` (as find_d_runs), else per test.
// ev gains the evaluations (one per primitive tested, the stacks' own count).
__device__ __forceinline__ int find_c_runs(CPTR(GeomDesc) G, int s, int e, const Ray32 &ray, int stacks, uint64_t &ev)
{
    CPTR(rf_primitive) P = cst(G->prims);
    if (!stacks || s >= e) {
        ev += (uint64_t)(e > s ? e - s : 0);
        return find_c(P, s, e, ray);
    }
    const Ray32 R1[1] = {ray};
    float tbest[1] = {finf()};
    int obj[1] = {-1};
    const int n_runs = G->n_c_runs;
    for (int j = cst(G->c_run_of)[s]; j < n_runs; ++j) {
        const CylRun run = ldc(cst(G->c_runs) + j);
        if (run.begin >= e) break;
        const int rb = max(run.begin, s), re = min(run.end, e);
        if (run.stack && re - rb >= 2) {
            uint32_t e1 = 0;
            stack_c<1>(P, cst(G->c_bands), run, rb, re, stacks >= 2 ? run.stack : 1, R1, tbest, obj, e1);
            ev += e1;
        } else {
            scan_c_range<1>(P, rb, re, R1, tbest, obj);
            ev += (uint64_t)(re - rb);
        }
    }
    return obj[0];
}

// BV body shared by compat and fp64 (MATH selects the arithmetic).  The
// final child scan of lanes whose winning bounding surface differs is done
// one distinct surface at a time: the leader's surface is read with
// readlane, so every scan runs over a wave-uniform range.  Child scans use
// the band stacks of the caller's order when `stacks` (cut to the child
// range; their flops are not counted apart: tests x the scene's mean F).
template <int MATH, typename RayT>
__device__ __forceinline__ int find_bv_cd(CPTR(GeomDesc) G, const RayT &ray, uint64_t &tests, int stacks)
{
    CPTR(rf_bounding_volume) B = cst(G->bvs);
    const int nbv = G->n_bvs;
    uint64_t sf = 0;
    // tests gain the child scans' evaluations
    const auto child = [&](int cs, int ce) {
        if constexpr (MATH == KM_COMPAT) return find_c_runs(G, cs, ce, ray, stacks, tests);
        else return find_d_runs(G, cs, ce, ray, stacks, sf, tests, true);
    };
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
                objectIndex = child(b.objStartIndex, b.objEndIndex);
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
            objectIndex = child(s, e);
            pending = false;
        }
    }
    return objectIndex;
}

#define ROW_S(P, i, R) __builtin_fmaf(P.r[4 * i], R.sx, __builtin_fmaf(P.r[4 * i + 1], R.sy, __builtin_fmaf(P.r[4 * i + 2], R.sz, P.r[4 * i + 3])))
#define ROW_C(P, i, R) __builtin_fmaf(P.r[4 * i], R.cx, __builtin_fmaf(P.r[4 * i + 1], R.cy, P.r[4 * i + 2] * R.cz))

template <int KIND, bool REFINE = true>
__device__ __forceinline__ float fast_test(const FastPrim &P, const Ray32 &R, bool &flag, float tau, float eps)
{
    if constexpr (KIND == SEG_CYL) {
        const float sx = ROW_S(P, 0, R), sy = ROW_S(P, 1, R), sz = ROW_S(P, 2, R);
        const float cx = ROW_C(P, 0, R), cy = ROW_C(P, 1, R), cz = ROW_C(P, 2, R);
        // b^2 - A C = A - (s x c)_z^2 (Lagrange): no C term, and the
        // rounding error scales with A rather than b^2
        const float A = __builtin_fmaf(cx, cx, cy * cy);
        const float b = __builtin_fmaf(sx, cx, sy * cy);
        const float w = __builtin_fmaf(sx, cy, -(sy * cx));
        const float D = __builtin_fmaf(-w, w, A);
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float tm = -b * rA;                       // tangent point
        const float t1 = __builtin_fmaf(-sq, rA, tm);
        const float t2 = __builtin_fmaf(sq, rA, tm);
        const float z1 = __builtin_fmaf(cz, t1, sz);
        const float z2 = __builtin_fmaf(cz, t2, sz);
        const bool e1 = t1 > eps, e2 = t2 > eps;
        const float a1 = __builtin_fabsf(z1), a2 = __builtin_fabsf(z2);
        if constexpr (!REFINE) {
            flag = false;
            return (e1 & (a1 <= 0.5f)) ? t1 : ((e2 & (a2 <= 0.5f)) ? t2 : finf());
        } else {
            // A root is certainly on the band for |z| < 1/2 - tau and
            // certainly off it for |z| > 1/2 + tau; in between it is a rim
            // flag.  A flagged pair's FP32 t is discarded, so the FP32
            // decision of an unflagged pair equals the |z| <= 1/2 test.
            const float lo = 0.5f - tau, hi = 0.5f + tau;
            const bool in1 = a1 < lo, in2 = a2 < lo;
            // near-tangent only where the tangent point lies on this band
            const float zm = __builtin_fmaf(cz, tm, sz);
            // in implies a <= hi (lo <= hi), so "a <= hi and not in" is an xor
            flag = ((__builtin_fabsf(D) <= tau * A) & (__builtin_fabsf(zm) <= hi)) |
                   (e1 & ((a1 <= hi) ^ in1)) | (e2 & ((a2 <= hi) ^ in2));
            return (e1 & in1) ? t1 : ((e2 & in2) ? t2 : finf());
        }
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
        const float tm = -b * rA;
        const float ta = __builtin_fmaf(-sq, rA, tm);
        const float tb = __builtin_fmaf(sq, rA, tm);
        const float t1 = __builtin_fminf(ta, tb);
        const float t2 = __builtin_fmaxf(ta, tb);
        const float w1 = __builtin_fmaf(cw, t1, sw) - P.k0;
        const float w2 = __builtin_fmaf(cw, t2, sw) - P.k0;
        const bool e1 = t1 > eps, e2 = t2 > eps;
        const float a1 = __builtin_fabsf(w1), a2 = __builtin_fabsf(w2);
        if constexpr (!REFINE) {
            flag = false;
            return (e1 & (a1 <= P.k1)) ? t1 : ((e2 & (a2 <= P.k1)) ? t2 : finf());
        } else {
            // band of half-width k1 with rims of 2 tau k1 (as the cylinder)
            const float lo = (1.0f - 2.0f * tau) * P.k1, hi = (1.0f + 2.0f * tau) * P.k1;
            const bool in1 = a1 < lo, in2 = a2 < lo;
            const float wm = __builtin_fmaf(cw, tm, sw) - P.k0;
            flag = ((__builtin_fabsf(D) <= tau * (b * b)) & (__builtin_fabsf(wm) <= hi)) |
                   (__builtin_fabsf(A) < tau * Axy) |
                   (e1 & ((a1 <= hi) ^ in1)) | (e2 & ((a2 <= hi) ^ in2));
            return (e1 & in1) ? t1 : ((e2 & in2) ? t2 : finf());
        }
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
        flag = (__builtin_fabsf(D) <= tau * (b * b)) & (b < 0.0f);
        return (t1 > eps) ? t1 : ((t2 > eps) ? t2 : finf());
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
            v = (t > eps) & (rr >= P.k0) & (rr <= P.k1);
        } else {
            v = (t > eps) & (__builtin_fabsf(hx) <= 1.0f) & (__builtin_fabsf(hy) <= 1.0f);
        }
        return v ? t : finf();
    }
}

template <bool REFINE = true>
__device__ __forceinline__ float fast_test_any(const FastPrim &P, const Ray32 &R, bool &flag, float tau, float eps)
{
    switch (P.type) {
    case RF_CYLINDER: return fast_test<SEG_CYL, REFINE>(P, R, flag, tau, eps);
    case RF_FRUSTUM: return fast_test<SEG_FRUSTUM, REFINE>(P, R, flag, tau, eps);
    case RF_SPHERE: return fast_test<SEG_SPHERE, REFINE>(P, R, flag, tau, eps);
    case RF_RECTANGLE: return fast_test<SEG_RECT, REFINE>(P, R, flag, tau, eps);
    default: return fast_test<SEG_RING, REFINE>(P, R, flag, tau, eps);
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
// list is re-solved in FP64 and merged into the nearest hit.  No FP64 code
// runs inside the primitive loop.  Once the list is full, flagged pairs are
// dropped and the first dropped key (position x 8 + ray slot) is kept in
// `resume`; the lane then re-scans the quadrics one ray slot at a time from
// there (pairs below `from` masked), re-solving as it goes, so every flagged
// pair is re-solved exactly once whatever the number of rays per thread: the
// result is the same for every launch shape.
constexpr int PEND_K = 4;
constexpr int NO_KEY = 0x7fffffff;

struct Pending {
    uint32_t *buf;     // LDS, PEND_K * blockDim entries: (index << 3) | ray slot
    int count;
    int resume;        // first dropped pair key, NO_KEY if none
    int from;          // re-scan: pairs with key < from are skipped
    float tau;         // relative flag margin (VGPR copy)
    float eps;         // EPS_F (VGPR copy)
    int slot0 = 0;     // band stacks: ray slot of the caller's first ray (a stack scanned in slices)
    int stack0 = -1;   // band stacks: list length to give back to on overflow (-1: at the call)
};

// Wave-wide minimum of a lane value (re-scan start).
__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return __builtin_amdgcn_readfirstlane(v);
}

// An integer zero the optimiser cannot see through: keeps a recomputation
// from being merged with the first computation (which would keep the first
// result's registers live across the FP64 stage).
__device__ __forceinline__ int opaque_zero()
{
    int z = 0;
    asm volatile("" : "+v"(z));
    return z;
}

// RESCAN: one ray (NR = 1) of lane slot `slot`, positions below pd.from
// skipped, `resume` records a position instead of a key.  act = false: the
// lane only rides along (its flags are dropped; its minimum is the caller's
// concern).
template <int KIND, int NR, bool REFINE, bool RESCAN>
__device__ __forceinline__ void seg_loop(CPTR(FastPrim) F, int begin, int end, const Ray32 (&R)[NR],
                                         float (&tb)[NR], int (&ib)[NR], Pending &pd, int slot = 0,
                                         bool act = true)
{
    if (begin >= end) return;
    // the next record's scalar load is issued a whole iteration ahead
    FastPrim next = ldc(F + begin);
    for (int w = begin; w < end; ++w) {
        const FastPrim q = vcopy(next);
        next = ldc(F + min(w + 1, end - 1));
        float t[NR];
        bool f[NR];
        bool any = false;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            t[i] = fast_test<KIND, REFINE>(q, R[i], f[i], pd.tau, pd.eps);
            if constexpr (RESCAN) {
                const bool skip = w < pd.from;
                f[i] = f[i] & !skip;
                t[i] = skip ? finf() : t[i];
            }
            f[i] = f[i] & act;
            any |= f[i];
        }
        if constexpr (REFINE && KIND <= SEG_SPHERE) {
            if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
                for (int i = 0; i < NR; ++i) {
                    if (f[i]) {
                        if (pd.count < PEND_K && pd.resume == NO_KEY) {
                            pd.buf[pd.count * blockDim.x + threadIdx.x] =
                                ((uint32_t)q.orig << 3) | (uint32_t)(RESCAN ? slot : i);
                            ++pd.count;
                        } else if (pd.resume == NO_KEY) {
                            pd.resume = RESCAN ? w : w * 8 + i;
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

// Band loop of stack_loop, bands e-1 down to b.  Wave-uniform variants:
//  AXIAL: the stack's z row has no x or y term (r[8] == r[9] == 0), so the
//    band's z origin is one FMA; the per-test scan's two extra FMAs add a
//    signed zero, which no |z| compare can tell apart;
//  TAN: some ray of the wave is near-tangent (finite TM), so the tangent-point
//    flag is evaluated; otherwise it is false for every ray;
//  ROOT1: some ray of the wave has a first root ahead of it (finite T1).
//    Otherwise root 1 is never on a band nor a rim flag (every ray starts
//    inside a wall's cylinder), and without the tangent flag a flagged pair
//    (a rim of root 2) can never also hold root 2 on the band (|z| < lo vs
//    lo <= |z|), so no pair needs excluding.
template <int NR, bool REFINE, bool AXIAL, bool TAN, bool ROOT1>
__device__ __forceinline__ void stack_bands(const FastPrim &h, CPTR(BandRec) BR, int b, int e, const Ray32 (&R)[NR],
                                            const float (&T1)[NR], const float (&T2)[NR], const float (&TM)[NR],
                                            const float (&CZ)[NR], int (&I1)[NR], int (&I2)[NR], float lo, float hi,
                                            Pending &pd)
{
    const int cnt0 = pd.stack0 >= 0 ? pd.stack0 : pd.count;
    // BR[-1] exists (the band array has a leading pad record), so the
    // prefetch of the next record needs no clamp
    CPTR(BandRec) p = BR + (e - 1);
    BandRec next = ldc(p);
    for (int w = e - 1; w >= b; --w) {
        float r11, of;
        vcopy2(next.r11, __int_as_float(next.orig), r11, of);
        const int orig = __float_as_int(of);
        --p;
        next = ldc(p);
        bool f[NR];
        bool any = false;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const float sz = AXIAL ? __builtin_fmaf(h.r[10], R[i].sz, r11)
                                   : __builtin_fmaf(h.r[8], R[i].sx,
                                                    __builtin_fmaf(h.r[9], R[i].sy, __builtin_fmaf(h.r[10], R[i].sz, r11)));
            const float a2 = __builtin_fabsf(__builtin_fmaf(CZ[i], T2[i], sz));
            const float a1 = ROOT1 ? __builtin_fabsf(__builtin_fmaf(CZ[i], T1[i], sz)) : 0.f;
            bool in1 = false, in2;
            if constexpr (REFINE) {
                // fast_test's terms: a root is on the band for |z| < lo and a
                // rim flag for lo <= |z| <= hi (its (|z| <= hi) ^ (|z| < lo);
                // med3 clamps |z| to [lo, hi], a NaN fails the equality)
                in2 = a2 < lo;
                f[i] = __builtin_amdgcn_fmed3f(a2, lo, hi) == a2;
                if constexpr (ROOT1) {
                    in1 = a1 < lo;
                    f[i] = f[i] | (__builtin_amdgcn_fmed3f(a1, lo, hi) == a1);
                }
                if constexpr (TAN) f[i] = f[i] | (__builtin_fabsf(__builtin_fmaf(CZ[i], TM[i], sz)) <= hi);
                any |= f[i];
                if constexpr (ROOT1 || TAN) {
                    in1 = in1 & !f[i];
                    in2 = in2 & !f[i];
                }
            } else {
                in2 = a2 <= 0.5f;
                if constexpr (ROOT1) in1 = a1 <= 0.5f;
            }
            if constexpr (ROOT1) I1[i] = in1 ? orig : I1[i];
            I2[i] = in2 ? orig : I2[i];
            // updates stay ahead of the branch below (sunk past it, the range
            // tests lose their |z| operand modifiers)
            asm volatile("" ::"v"(I1[i]), "v"(I2[i]));
        }
        if (REFINE && __builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                if (f[i]) {
                    if (pd.count < PEND_K && pd.resume == NO_KEY) {
                        pd.buf[pd.count * blockDim.x + threadIdx.x] = ((uint32_t)orig << 3) | (uint32_t)(pd.slot0 + i);
                        ++pd.count;
                    } else if (pd.resume == NO_KEY) {
                        pd.resume = b * 8;
                        pd.count = cnt0;
                    }
                }
            }
        }
    }
}

template <int NR, bool REFINE, bool AXIAL>
__device__ __forceinline__ void stack_bands_sel(bool tan, bool root1, const FastPrim &h, CPTR(BandRec) BR, int b,
                                                int e, const Ray32 (&R)[NR], const float (&T1)[NR],
                                                const float (&T2)[NR], const float (&TM)[NR], const float (&CZ)[NR],
                                                int (&I1)[NR], int (&I2)[NR], float lo, float hi, Pending &pd)
{
    if (tan) {
        if (root1) stack_bands<NR, REFINE, AXIAL, true, true>(h, BR, b, e, R, T1, T2, TM, CZ, I1, I2, lo, hi, pd);
        else stack_bands<NR, REFINE, AXIAL, true, false>(h, BR, b, e, R, T1, T2, TM, CZ, I1, I2, lo, hi, pd);
    } else {
        if (root1) stack_bands<NR, REFINE, AXIAL, false, true>(h, BR, b, e, R, T1, T2, TM, CZ, I1, I2, lo, hi, pd);
        else stack_bands<NR, REFINE, AXIAL, false, false>(h, BR, b, e, R, T1, T2, TM, CZ, I1, I2, lo, hi, pd);
    }
}

// z origin of band record r11 of a stack with head h (fast_test's z row).
template <bool AXIAL>
__device__ __forceinline__ float band_sz(const FastPrim &h, float r11, const Ray32 &R)
{
    return AXIAL ? __builtin_fmaf(h.r[10], R.sz, r11)
                 : __builtin_fmaf(h.r[8], R.sx, __builtin_fmaf(h.r[9], R.sy, __builtin_fmaf(h.r[10], R.sz, r11)));
}

// Band records of a searched stack, read either through the constant address
// space (global memory; the probes' addresses differ per lane, so these are
// vector loads served by the caches) or from the unit's copy in LDS, whose
// lower latency shortens the dependent guess -> check -> walk chain.
struct BandsGlobal {
    CPTR(BandRec) p;
    __device__ __forceinline__ BandRec operator[](int i) const { return ldc(p + i); }
};

#define LDS_PTR(T) const __attribute__((address_space(3))) T *

// LDS copy of the records of cylinder positions [base, base + n): one 64-bit
// read per record.
struct BandsLds {
    LDS_PTR(uint64_t) p;
    int base;
    __device__ __forceinline__ BandRec operator[](int i) const
    {
        const uint64_t v = p[i - base];
        return BandRec{__uint_as_float((uint32_t)v), (int32_t)(v >> 32)};
    }
};

// Band stack whose band offsets r[11] are sorted over the positions
// (ASC: non-decreasing, else non-increasing; all finite).  Each z the band
// loop computes, fma(cz, t, sz(w)) with sz(w) = fma(..., r[11](w)), is then a
// monotone function of the position (an FMA is non-decreasing in its addend),
// so the positions where a root's |z| <= hi -- the only ones where fast_test's
// on-band or rim terms can hold -- form one run, found by a binary search for
// its first position and a walk to its end.  The band loop's per-band work is
// done on the positions of the runs of root 2, root 1 and the tangent point
// (each position once), so the first band holding each root, the flagged
// pairs and the list overflow are the band loop's.  I1 / I2 return positions
// (e: none).  With COUNT, ev gains the z evaluations (probes and walked
// positions) of the ray slots whose bit is set in vmask.  BR: BandsGlobal or
// BandsLds.
template <int NR, bool REFINE, bool AXIAL, bool ASC, bool COUNT, class Bands>
__device__ __forceinline__ void stack_search(const FastPrim &h, const Bands &BR, int b, int e, float r0,
                                             float inv, const Ray32 (&R)[NR], const float (&T1)[NR], const float (&T2)[NR],
                                             const float (&TM)[NR], const float (&CZ)[NR], int (&I1)[NR],
                                             int (&I2)[NR], float lo, float hi, bool wroot1, bool wtan, Pending &pd,
                                             uint32_t &ev, uint32_t vmask)
{
    const int cnt0 = pd.stack0 >= 0 ? pd.stack0 : pd.count;
    // the band loop's per-band work on position w of ray slot i
    const auto band = [&](int i, int w, int orig, float sz) {
        const float a1 = __builtin_fabsf(__builtin_fmaf(CZ[i], T1[i], sz));
        const float a2 = __builtin_fabsf(__builtin_fmaf(CZ[i], T2[i], sz));
        bool in1, in2, f = false;
        if constexpr (REFINE) {
            in1 = a1 < lo;
            in2 = a2 < lo;
            f = (__builtin_amdgcn_fmed3f(a1, lo, hi) == a1) | (__builtin_amdgcn_fmed3f(a2, lo, hi) == a2) |
                (__builtin_fabsf(__builtin_fmaf(CZ[i], TM[i], sz)) <= hi);
            in1 = in1 & !f;
            in2 = in2 & !f;
        } else {
            in1 = a1 <= 0.5f;
            in2 = a2 <= 0.5f;
        }
        I1[i] = (in1 && w < I1[i]) ? w : I1[i];
        I2[i] = (in2 && w < I2[i]) ? w : I2[i];
        if (REFINE && f) {
            if (pd.count < PEND_K && pd.resume == NO_KEY) {
                pd.buf[pd.count * blockDim.x + threadIdx.x] = ((uint32_t)orig << 3) | (uint32_t)(pd.slot0 + i);
                ++pd.count;
            } else if (pd.resume == NO_KEY) {
                pd.resume = b * 8;
                pd.count = cnt0;
            }
        }
    };
    // run of root K (2, 1, tangent point) for every ray slot: first position
    // by search, then the walk while |z| <= hi.  A position is in the run of
    // an earlier root exactly when that root's |z| <= hi there (monotone z),
    // so those are skipped.
    const auto run = [&](const int K) {
        const auto root = [&](int i) { return K == 2 ? T2[i] : (K == 1 ? T1[i] : TM[i]); };
        const auto zof = [&](int i, float r11) {
            return __builtin_fmaf(CZ[i], root(i), band_sz<AXIAL>(h, r11, R[i]));
        };
        const auto pass = [&](float z) { return ASC ? z >= -hi : z <= hi; };
        // first guess, z(w) being about fma(cz, t, sz with r[11] = 0) + r[11](w);
        // the records before it, at it and after it are loaded together: the
        // check of the guess needs the first two, and the walk usually ends
        // at the third
        int g[NR];
        BandRec rg[NR], rn[NR];
        bool ok[NR];
        bool off = false;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const float k0 = __builtin_fmaf(CZ[i], root(i), band_sz<AXIAL>(h, 0.0f, R[i]));
            const int gi = band_guess(b, e, r0, inv, ASC ? -hi - k0 : hi - k0);
            const BandRec rm = BR[max(gi - 1, b)];
            rg[i] = BR[min(gi, e - 1)];
            rn[i] = BR[min(gi + 1, e - 1)];
            if constexpr (COUNT) ev += (vmask >> i) & 1u ? 2u : 0u;
            ok[i] = (gi >= e || pass(zof(i, rg[i].r11))) && (gi <= b || !pass(zof(i, rm.r11)));
            g[i] = gi;
            off |= !ok[i];
        }
        if (__builtin_amdgcn_ballot_w64(off)) {
            int bs[NR];
            band_lower_bound<NR>(b, e, [&](int w, int i) {
                if constexpr (COUNT) ev += (vmask >> i) & 1u;
                return pass(zof(i, BR[w].r11));
            }, bs);
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                if (!ok[i]) {
                    g[i] = bs[i];
                    rg[i] = BR[min(bs[i], e - 1)];
                    rn[i] = BR[min(bs[i] + 1, e - 1)];
                }
            }
        }
        // the walk; its first two positions use the records already loaded
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            int w = g[i];
            bool go = w < e;
            const auto step = [&](const BandRec &br) {
                const float sz = band_sz<AXIAL>(h, br.r11, R[i]);
                if constexpr (COUNT) ev += (vmask >> i) & 1u;
                if (!(__builtin_fabsf(__builtin_fmaf(CZ[i], root(i), sz)) <= hi)) {
                    go = false;
                    return;
                }
                bool done = false;
                if (K != 2) done = __builtin_fabsf(__builtin_fmaf(CZ[i], T2[i], sz)) <= hi;
                if (K == 0) done = done || __builtin_fabsf(__builtin_fmaf(CZ[i], T1[i], sz)) <= hi;
                if (!done) band(i, w, br.orig, sz);
                ++w;
                go = w < e;
            };
            if (go) step(rg[i]);
            if (go) step(rn[i]);
            while (go) step(BR[w]);
        }
    };
    run(2);
    if (wroot1) run(1);
    if (REFINE && wtan) run(0);
}

template <int NR, bool REFINE, bool AXIAL, bool COUNT, class Bands>
__device__ __forceinline__ void stack_search_sel(bool asc, const FastPrim &h, const Bands &BR, int b, int e,
                                                 float r0, float inv, const Ray32 (&R)[NR], const float (&T1)[NR],
                                                 const float (&T2)[NR], const float (&TM)[NR],
                                                 const float (&CZ)[NR], int (&I1)[NR], int (&I2)[NR], float lo,
                                                 float hi, bool wroot1, bool wtan, Pending &pd, uint32_t &ev,
                                                 uint32_t vmask)
{
    if (asc)
        stack_search<NR, REFINE, AXIAL, true, COUNT>(h, BR, b, e, r0, inv, R, T1, T2, TM, CZ, I1, I2, lo, hi, wroot1, wtan,
                                                     pd, ev, vmask);
    else
        stack_search<NR, REFINE, AXIAL, false, COUNT>(h, BR, b, e, r0, inv, R, T1, T2, TM, CZ, I1, I2, lo, hi, wroot1, wtan,
                                                      pd, ev, vmask);
}

// Band stack [b, e) of the cylinder segment (CylRun.stack >= 1): the records
// are bitwise equal except the z offset r[11], so each ray's x-y quadric, its
// roots and its z direction are computed once, with the operations of
// fast_test<SEG_CYL>, and each band then needs only its z origin row, z at
// both roots (and at the tangent point) and the range compares.  Every
// per-band decision is fast_test's on that band's own record, bit for bit:
//  - a root failing t > eps becomes NaN, so every compare on its z is false
//    (fast_test's e1 / e2 terms); the tangent point likewise when |D| > tau A;
//  - fast_test's candidate on a band is t1 if z1 is on it, else t2 if z2 is,
//    and t1 <= t2 (sqrt >= 0, 1/A >= 0; t1 == t2 makes z1 == z2), so the
//    nearest hit of the stack is t1 at the first band holding z1, else t2 at
//    the first band holding z2 -- the bands are walked last to first and the
//    last band seen holding each root is kept -- and it enters the running
//    minimum with the strict < of the per-test scan;
//  - fast_test's flag terms are evaluated on the same values; its rim term
//    (|z| <= hi) ^ (|z| < lo) is lo <= |z| <= hi, here med3(|z|, lo, hi) == |z|
//    (one vector compare instead of a compare and a scalar xor).
// A flagged pair is queued as in seg_loop.  If the lane's list fills inside the
// stack, the entries it queued from this stack are given back and the re-scan
// starts at the stack's first band (resume = b x 8), so every flagged pair is
// still re-solved exactly once.
// mode 1 walks every band; 2 and 3 (a stack sorted by r[11], non-decreasing or
// non-increasing) search the bands instead (stack_search) at 1 or 2 rays per
// thread.  At 4 and 8 the search's registers would spill in the 96 (4) or
// fill the 192 (8) a wave has at that shape's occupancy, so those shapes walk
// every band; the autotuner weighs the two.
// fl gains the algorithmic flops of the band work the wave ran for the ray
// slots whose bit is set in vmask, ev the evaluations: one per (ray, stack)
// for the shared quadric and one per band whose z range is checked (the
// search's probes and walked positions, or every band of a walked stack).
// Without COUNT nothing is counted (fl and ev are left alone: the launch
// reports no counts).  BL: the search reads the band records from
// this LDS copy when it is set (positions [BL.base, BL.base + n)).
//
// Analytic positions (run.an_margin > 0, axial z row; CylRun): before the
// search, each ray slot's roots are placed in the progression of band offsets
// in FP64.  With Z* = c_z t + r[10] s_z + r[11] the exact z of a root on a
// band and zof its FP32 evaluation (two roundings), |zof - Z*| <= E =
// 2^-24 (|c_z t| + 4.01) on every band where |Z*| <= 2.  Let dc be the distance
// of x from the nearest position kc in [0, n) and delta the model spacing
// (|delta| >= 1), dev the offsets' deviation from it, and
// m = tau + dev + 2E + (FP64 and compare rounding).  Then
//  - |delta| dc <= 1/2 - m: band kc has |zof| < 1/2 - tau (on it, no rim flag)
//    and every other band |zof| > 1/2 + tau (off it, no flag);
//  - |delta| dc >= 1/2 + m: every band has |zof| > 1/2 + tau;
// so the searched bands would decide exactly this.  A slot whose roots are all
// so decided (and whose tangent point is not flagged) enters the running
// minimum at once and leaves the search (its roots become NaN there); the
// search runs, for the other slots only, when any lane has one.
template <int NR, bool REFINE, bool COUNT = true, bool BT = false>
__device__ __forceinline__ void stack_loop(CPTR(FastPrim) F, CPTR(BandRec) BR, const CylRun &run, int mode,
                                           const Ray32 (&R)[NR], float (&tb)[NR], int (&ib)[NR], Pending &pd,
                                           uint32_t &fl, uint32_t &ev, uint32_t vmask, BandsLds BL = BandsLds{})
{
    const int b = run.begin, e = run.end;
    const float inv = run.inv;
    const FastPrim hs = ldc(F + b);
    const FastPrim h = vcopy(hs);
    // wave-uniform choice of the band loop: integer tests on the scalar copy (+-0)
    const bool axial = ((__float_as_uint(hs.r[8]) | __float_as_uint(hs.r[9])) << 1) == 0u;
    const float tau = pd.tau, eps = pd.eps;
    const float lo = 0.5f - tau, hi = 0.5f + tau;
    const float qnan = __builtin_nanf("");
    float T1[NR], T2[NR], TM[NR], CZ[NR];
    int I1[NR], I2[NR];
    bool tan = false, root1 = false;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const float sx = ROW_S(h, 0, R[i]), sy = ROW_S(h, 1, R[i]);
        const float cx = ROW_C(h, 0, R[i]), cy = ROW_C(h, 1, R[i]);
        CZ[i] = ROW_C(h, 2, R[i]);
        const float A = __builtin_fmaf(cx, cx, cy * cy);
        const float bq = __builtin_fmaf(sx, cx, sy * cy);
        const float w = __builtin_fmaf(sx, cy, -(sy * cx));
        const float D = __builtin_fmaf(-w, w, A);
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float tm = -bq * rA;
        const float t1 = __builtin_fmaf(-sq, rA, tm);
        const float t2 = __builtin_fmaf(sq, rA, tm);
        T1[i] = t1 > eps ? t1 : qnan;
        T2[i] = t2 > eps ? t2 : qnan;
        root1 |= t1 > eps;
        const bool near = REFINE && __builtin_fabsf(D) <= tau * A;
        TM[i] = near ? tm : qnan;
        tan |= near;
        // materialised once: keeps the selects out of the band loop
        asm volatile("" : "+v"(T1[i]), "+v"(T2[i]), "+v"(TM[i]), "+v"(CZ[i]));
        I1[i] = -1;
        I2[i] = -1;
    }
    const bool wtan = REFINE && __builtin_amdgcn_ballot_w64(tan);
    bool wroot1 = __builtin_amdgcn_ballot_w64(root1);
    if (NR <= 2 && mode >= 2) {
        // positions, e = none
#pragma unroll
        for (int i = 0; i < NR; ++i) { I1[i] = e; I2[i] = e; }
        // z evaluations of this stack: the only counter inside the search
        uint32_t zs = 0;
        const float r0 = run.r0;
        const auto orig_at = [&](int k) {
            if (run.ostep) return run.orig0 + (k - b) * run.ostep;
            return BT ? BL[k].orig : ldc(BR + k).orig;
        };
        uint32_t smask = 0;       // ray slots placed analytically
        bool rest = true;         // some lane has a slot left for the search
        if (axial && run.an_margin > 0.0f) {
            const float mfix = tau + run.an_margin + 0x1.1p-21f;   // 2^-23 x 4.25
            const float ha = 0.5f * run.an_ainv;
            const double nm1 = (double)(e - b - 1);
            bool left = false;
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                const double P = __builtin_fma((double)h.r[10], (double)R[i].sz, (double)r0);
                const double czd = (double)CZ[i];
                // position of root t (e: none); sure = the decision is certain
                const auto place = [&](float t, bool &sure) {
                    const double x = __builtin_fma(czd, (double)t, P) * run.an_ninv;
                    const double k = __builtin_fmin(__builtin_fmax(__builtin_rint(x), 0.0), nm1);
                    const float dc = (float)__builtin_fabs(x - k);
                    const float m = __builtin_fmaf(0x1.002p-23f, __builtin_fabsf(CZ[i] * t), mfix);
                    const bool on = dc <= __builtin_fmaf(-m, run.an_ainv, ha);
                    sure = on || !(dc < __builtin_fmaf(m, run.an_ainv, ha));
                    return on ? b + (int)k : e;
                };
                bool s2, s1 = true;
                const int k2 = place(T2[i], s2);
                int k1 = e;
                if (wroot1) k1 = place(T1[i], s1);
                const bool sure = s1 && s2 && !(TM[i] == TM[i]);
                if (sure) {
                    const bool h1 = k1 < e;
                    const float c = h1 ? T1[i] : (k2 < e ? T2[i] : finf());
                    if (c < tb[i]) { tb[i] = c; ib[i] = orig_at(h1 ? k1 : k2); }
                    T1[i] = __builtin_nanf("");
                    T2[i] = __builtin_nanf("");
                    smask |= 1u << i;
                }
                left |= !sure;
            }
            if constexpr (COUNT) {
                // one position per root per slot: the z origin row, z at the root and the scale
                const uint32_t q = (uint32_t)__builtin_popcount(vmask & smask) * (wroot1 ? 2u : 1u);
                fl += q * (kFlopsBandZ0Axial + kFlopsBandRoot + 1u);
                ev += q;
            }
            rest = __builtin_amdgcn_ballot_w64(left);
            if (rest) {
                bool r1 = false;
#pragma unroll
                for (int i = 0; i < NR; ++i) r1 |= T1[i] == T1[i];
                wroot1 = __builtin_amdgcn_ballot_w64(r1);
            }
        }
        if (!rest) {
            if constexpr (COUNT) {
                const uint32_t q = (uint32_t)__builtin_popcount(vmask);
                fl += q * kFlopsStack;
                ev += q;
            }
            return;
        }
        const uint32_t xmask = vmask & ~smask;
        const auto search = [&](const auto &src) {
            if (axial)
                stack_search_sel<NR, REFINE, true, COUNT>(mode == 2, h, src, b, e, r0, inv, R, T1, T2, TM, CZ, I1, I2, lo,
                                                          hi, wroot1, wtan, pd, zs, xmask);
            else
                stack_search_sel<NR, REFINE, false, COUNT>(mode == 2, h, src, b, e, r0, inv, R, T1, T2, TM, CZ, I1, I2, lo,
                                                           hi, wroot1, wtan, pd, zs, xmask);
        };
        if constexpr (BT) search(BL);
        else search(BandsGlobal{BR});
        if constexpr (COUNT) {
            const uint32_t q = (uint32_t)__builtin_popcount(vmask);
            fl += q * kFlopsStack + zs * ((axial ? kFlopsBandZ0Axial : kFlopsBandZ0) + kFlopsBandRoot);
            ev += q + zs;
        }
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const bool h1 = I1[i] < e;
            const float c = h1 ? T1[i] : (I2[i] < e ? T2[i] : finf());
            const int k = min(h1 ? I1[i] : I2[i], e - 1);
            if (c < tb[i]) { tb[i] = c; ib[i] = BT ? BL[k].orig : BR[k].orig; }
        }
        return;
    }
    if (axial) stack_bands_sel<NR, REFINE, true>(wtan, wroot1, h, BR, b, e, R, T1, T2, TM, CZ, I1, I2, lo, hi, pd);
    else stack_bands_sel<NR, REFINE, false>(wtan, wroot1, h, BR, b, e, R, T1, T2, TM, CZ, I1, I2, lo, hi, pd);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const bool h1 = I1[i] >= 0;
        const float c = h1 ? T1[i] : (I2[i] >= 0 ? T2[i] : finf());
        const int ci = h1 ? I1[i] : I2[i];
        if (c < tb[i]) { tb[i] = c; ib[i] = ci; }
    }
    if constexpr (COUNT) {
        const uint32_t sf = kFlopsStack + (uint32_t)(e - b) * ((axial ? kFlopsBandZ0Axial : kFlopsBandZ0) +
                                                               (wroot1 ? 2u : 1u) * kFlopsBandRoot);
        fl += (uint32_t)__builtin_popcount(vmask) * sf;
        ev += (uint32_t)__builtin_popcount(vmask) * (uint32_t)(1 + e - b);
    }
}

// FP64 stage of the deferred re-solves (see Pending).  The flagged ray is
// regenerated from its index (emission is a pure function of the random
// block), so the scan's ray registers are dead here.
// ONE: every entry belongs to one ray, merged into tb[0] / ib[0].
// ray_of(k): unit-local index of the ray in slot k.
template <int NR, bool ONE = false, class RayOf>
__device__ __forceinline__ void resolve_pending(const KParams &kp, CPTR(GeomDesc) G, const Emitter &E,
                                                uint32_t gid, int emitter, int64_t wi0, RayOf ray_of,
                                                float (&tb)[NR], int (&ib)[NR], Pending &pd, uint32_t &nref)
{
    CPTR(rf_primitive) P = cst(G->prims);
    for (int e = 0; e < pd.count; ++e) {
        const uint32_t ent = pd.buf[e * blockDim.x + threadIdx.x];
        const int prim = (int)(ent >> 3);
        const int k = (int)(ent & 7u);
        const int64_t rl = ray_of(k);
        float f[4], r4[4];
        workitem_floats(kp, gid, emitter, wi0 + (rl >> 2), f);
        rotate_rands(f, (int)(rl & 3), r4);
        const Ray64 r64 = widen(emit_fast(&E, r4));
        const rf_primitive p = ldc(P + prim);
        const float t = (float)hit_d_quadric(p.type, p.invm, p.util1, r64);
#pragma unroll
        for (int i = 0; i < NR; ++i)
            if ((ONE || k == i) && (t < tb[i] || (t == tb[i] && prim < ib[i]))) { tb[i] = t; ib[i] = prim; }
    }
    nref += (uint32_t)pd.count;
    pd.count = 0;
}

// Pairs dropped from a full list (first dropped key pd.resume): re-scan the
// quadrics for one ray slot at a time, from the first position whose pair of
// that slot may have been dropped, re-solving each time the list fills.
// Segment s of the scanned record array F spans [sb[s], se[s]).
template <int NR, class RayOf>
__device__ __forceinline__ void finish_dropped(const KParams &kp, CPTR(GeomDesc) G, const Emitter &E,
                                               uint32_t gid, int emitter, int64_t wi0, RayOf ray_of,
                                               CPTR(FastPrim) F, CPTR(int32_t) sb, CPTR(int32_t) se,
                                               float (&tb)[NR], int (&ib)[NR], Pending &pd, uint32_t &nref,
                                               uint32_t &nover)
{
    const int K = pd.resume;
#pragma unroll 1
    for (int i = 0; i < NR; ++i) {
        int from = K == NO_KEY ? NO_KEY : (K >> 3) + (i < (K & 7) ? 1 : 0);
        float tr[1] = {finf()};
        int ir[1] = {-1};
        while (__builtin_amdgcn_ballot_w64(from != NO_KEY)) {
            nover += from != NO_KEY;
            Ray32 R1[1];
            {
                float f[4], r4[4];
                const int64_t rl = ray_of(i) + opaque_zero();
                workitem_floats(kp, gid, emitter, wi0 + (rl >> 2), f);
                rotate_rands(f, (int)(rl & 3), r4);
                R1[0] = emit_fast(&E, r4);
            }
            pd.count = 0;
            pd.resume = NO_KEY;
            pd.from = from;
            const int lo = wave_min(from);
            const auto beg = [&](int sg) { return sb[sg] > lo ? sb[sg] : lo; };
            seg_loop<SEG_CYL, 1, true, true>(F, beg(SEG_CYL), se[SEG_CYL], R1, tr, ir, pd, i);
            seg_loop<SEG_FRUSTUM, 1, true, true>(F, beg(SEG_FRUSTUM), se[SEG_FRUSTUM], R1, tr, ir, pd, i);
            seg_loop<SEG_SPHERE, 1, true, true>(F, beg(SEG_SPHERE), se[SEG_SPHERE], R1, tr, ir, pd, i);
            if (__builtin_amdgcn_ballot_w64(pd.count > 0))
                resolve_pending<1, true>(kp, G, E, gid, emitter, wi0, ray_of, tr, ir, pd, nref);
            from = pd.resume;
        }
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (k == i && (tr[0] < tb[k] || (tr[0] == tb[k] && ir[0] < ib[k]))) { tb[k] = tr[0]; ib[k] = ir[0]; }
    }
    pd.resume = NO_KEY;
}

// Rings (discs, annuli) whose plane is z = z0 in world coordinates and whose
// object frame is a similarity of world x-y (records [begin, end) of the ring
// segment, re-encoded on the host: r[0] = z0, r[1], r[2] = the centre's world
// x, y, k0, k1 = the r^2 range in world units).  The hit is tested in world
// coordinates: t = (z0 - s_z) / c_z with 1/c_z computed once per ray for all
// of them, and r^2 = (s_x + t c_x - x0)^2 + (s_y + t c_y - y0)^2.  No FP32
// conditioning flag, as for every plane.
template <int NR>
__device__ __forceinline__ void zring_loop(CPTR(FastPrim) F, int begin, int end, const Ray32 (&R)[NR],
                                           float (&tb)[NR], int (&ib)[NR], float eps)
{
    if (begin >= end) return;
    float rcz[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) rcz[i] = rcp_fast(R[i].cz);
    FastPrim next = ldc(F + begin);
    for (int w = begin; w < end; ++w) {
        // the five fields the test reads, in VGPRs (three moves)
        float z0, x0, y0, k0, k1;
        vcopy2(next.r[0], next.r[1], z0, x0);
        y0 = vcopy(next.r[2]);
        vcopy2(next.k0, next.k1, k0, k1);
        const int orig = next.orig;
        next = ldc(F + min(w + 1, end - 1));
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const float t = (z0 - R[i].sz) * rcz[i];
            const float dx = __builtin_fmaf(R[i].cx, t, R[i].sx) - x0;
            const float dy = __builtin_fmaf(R[i].cy, t, R[i].sy) - y0;
            const float rr = __builtin_fmaf(dx, dx, dy * dy);
            const float c = ((t > eps) & (rr >= k0) & (rr <= k1)) ? t : finf();
            if (c < tb[i]) { tb[i] = c; ib[i] = orig; }
        }
    }
}

// ---------------------------------------------------------------------------
// The hierarchy over a scene's per-test primitives (BvhNode), leaves typed
// ---------------------------------------------------------------------------
#define GPTR(T) const __attribute__((address_space(1))) T *

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Copy of a record read through the global address space with 16-byte vector
// loads (per-lane addresses: a leaf's records and the nodes a lane visits).
template <typename T>
__device__ __forceinline__ T ldg(GPTR(T) p)
{
    static_assert(sizeof(T) % 16 == 0, "record must be 16-byte sized");
    u32x4 d[sizeof(T) / 16];
    GPTR(u32x4) q = (GPTR(u32x4))p;
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 16; ++i) d[i] = q[i];
    T v;
    __builtin_memcpy(&v, d, sizeof(T));
    return v;
}

// Slab terms of one ray in the scene's box frame (world x = bvh_c + bvh_s h):
// with i = 1/c per axis (|c| below 1e-30 taken as +-1e-30, so no 0 x inf
// arises), inv = bvh_s i (exact: bvh_s is a power of two) and
// ns = (bvh_c - s) i; the ray crosses the plane h_a = w at
// t = fma(w, inv[a], ns[a]).
struct BoxRay {
    float inv[3], ns[3];
};

__device__ __forceinline__ BoxRay box_ray(const Ray32 &R, CPTR(GeomDesc) G)
{
    BoxRay b;
    const float c[3] = {R.cx, R.cy, R.cz}, s[3] = {R.sx, R.sy, R.sz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ca = __builtin_fabsf(c[a]) < 1e-30f ? __builtin_copysignf(1e-30f, c[a]) : c[a];
        const float i = rcp_fast(ca);
        b.ns[a] = (G->bvh_c[a] - s[a]) * i;
        b.inv[a] = G->bvh_s * i;
    }
    return b;
}

// Child k's reference of a node: the index (sign-extended; -1: no child) and
// the leaf word count | segment << 3 (0: not a leaf).
__device__ __forceinline__ int bvh_index(uint32_t r)
{
    return (int32_t)(r << (32 - BVH_INDEX_BITS)) >> (32 - BVH_INDEX_BITS);
}
__device__ __forceinline__ int bvh_leafword(uint32_t r) { return (int)(r >> BVH_INDEX_BITS); }

__device__ __forceinline__ float box_half(const BvhNode &nd, int i)
{
    return (float)__builtin_bit_cast(_Float16, nd.box[i]);
}

// Entry and exit distances of the ray's segment [0, tmax] through child k's
// box of node nd; the segment crosses the box when enter <= leave.
__device__ __forceinline__ void box_span(const BoxRay &b, const BvhNode &nd, int k, float tmax, float &enter,
                                         float &leave)
{
    float n[3], f[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = __builtin_fmaf(box_half(nd, 6 * k + a), b.inv[a], b.ns[a]);
        const float v = __builtin_fmaf(box_half(nd, 6 * k + 3 + a), b.inv[a], b.ns[a]);
        n[a] = __builtin_fminf(u, v);
        f[a] = __builtin_fmaxf(u, v);
    }
    enter = __builtin_fmaxf(__builtin_fmaxf(n[0], n[1]), __builtin_fmaxf(n[2], 0.0f));
    leave = __builtin_fminf(__builtin_fminf(f[0], f[1]), __builtin_fminf(f[2], tmax));
}

template <int KIND>
__device__ __forceinline__ constexpr uint32_t kind_flops()
{
    return (uint32_t)flops_per_test(KIND == SEG_CYL ? RF_CYLINDER : KIND == SEG_FRUSTUM ? RF_FRUSTUM
                                    : KIND == SEG_SPHERE ? RF_SPHERE : KIND == SEG_RING ? RF_DISC : RF_RECTANGLE);
}

// The n records of one leaf from c, of type segment KIND (q0: the first,
// loaded by the caller): tested as seg_loop tests them.  Each record carries its traversal position p (FastPrim::type): a
// hit at the best distance replaces the best only if that was found earlier
// in this walk at a larger position (bp), so the walk's result is the linear
// scan's (strict <, the earlier position wins a tie; positions increase from
// segment to segment) whatever the visit order.  Flagged pairs queue as in
// seg_loop; once the list is full the rest are dropped and `resume` falls to
// (the cylinder segment's first position, slot 0) if it is above, so every
// slot re-scans the quadric segments in full (finish_dropped).  The slots
// walk one after the other, so a later slot can drop a pair that precedes
// the first drop in scan order: hence the minimum.
template <int KIND, bool REFINE, bool COUNT>
__device__ __forceinline__ void bvh_leaf(GPTR(FastPrim) L, const FastPrim &q0, int c, int n, const Ray32 &R, float &tb,
                                         int &ib, int &bp, Pending &pd, int slot, int key0, uint32_t &fl,
                                         uint32_t &ev)
{
    for (int j = 0; j < n; ++j) {
        const FastPrim q = j == 0 ? q0 : ldg(L + c + j);
        bool f;
        float t = fast_test<KIND, REFINE>(q, R, f, pd.tau, pd.eps);
        if constexpr (REFINE && KIND <= SEG_SPHERE) {
            if (f) {
                if (pd.count < PEND_K && pd.resume == NO_KEY) {
                    pd.buf[pd.count * blockDim.x + threadIdx.x] = ((uint32_t)q.orig << 3) | (uint32_t)slot;
                    ++pd.count;
                } else {
                    pd.resume = min(pd.resume, key0);
                }
                t = finf();
            }
        }
        if (t < tb || (t == tb && q.type < bp)) { tb = t; ib = q.orig; bp = q.type; }
    }
    if constexpr (COUNT) {
        fl += (uint32_t)n * kind_flops<KIND>();
        ev += (uint32_t)n;
    }
}

// One node visit of a hierarchy walk (nd: the node's record, loaded): the
// node's two child boxes against the ray's segment [0, tb], `node` set to the
// next node to visit (the nearer crossed inner child, the other one pushed on
// the lane's LDS stack, entries blockDim.x apart; else the top of the stack;
// -1 when the walk is over) and its record loaded into nd, then the
// primitives of the crossed leaf children tested.  The visiting order and the
// tb each box is tested against are those of testing the leaves first: a
// node's boxes are tested only when it is visited, after the leaves before
// it.  A pruned box holds no hit at or before tb (the host pads every box
// beyond the FP32 error of hits and of slab distances, and rounds its stored
// corners outward), so neither an FP32 hit nor an FP64 re-solve inside it
// could change the result.
template <bool REFINE, bool COUNT>
__device__ __forceinline__ void bvh_step(GPTR(BvhNode) N, GPTR(FastPrim) L, int &node, BvhNode &nd, int &sp, int &bp,
                                         int key0, const Ray32 &R, const BoxRay &B, float &tb, int &ib, Pending &pd,
                                         int slot, uint32_t *stk, uint32_t &fl, uint32_t &ev)
{
    float e0, l0, e1, l1;
    box_span(B, nd, 0, tb, e0, l0);
    box_span(B, nd, 1, tb, e1, l1);
    const int c[2] = {bvh_index(nd.ref[0]), bvh_index(nd.ref[1])};
    const int n[2] = {bvh_leafword(nd.ref[0]), bvh_leafword(nd.ref[1])};
    const bool h0 = e0 <= l0;
    const bool h1 = (e1 <= l1) & (c[1] >= 0);
    if constexpr (COUNT) {
        fl += 2u * kFlopsBox;
        ev += 2u;
    }
    const bool leaf[2] = {h0 && n[0] > 0, h1 && n[1] > 0};
    // the crossed leaves' first records, then the next node: their loads
    // are in flight together, and the next node's under the leaf tests
    FastPrim q[2];
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (leaf[k]) q[k] = ldg(L + c[k]);
    const bool i0 = h0 & (n[0] == 0), i1 = h1 & (n[1] == 0);
    if (i0 & i1) {
        const bool near0 = e0 <= e1;
        stk[sp * blockDim.x] = (uint32_t)(near0 ? c[1] : c[0]);
        ++sp;
        node = near0 ? c[0] : c[1];
    } else if (i0) {
        node = c[0];
    } else if (i1) {
        node = c[1];
    } else {
        node = sp > 0 ? (int)stk[--sp * blockDim.x] : -1;
    }
    if (node >= 0) nd = ldg(N + node);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (leaf[k]) {
            const int cnt = n[k] & 7;
            switch (n[k] >> 3) {
            case SEG_CYL: bvh_leaf<SEG_CYL, REFINE, COUNT>(L, q[k], c[k], cnt, R, tb, ib, bp, pd, slot, key0, fl, ev); break;
            case SEG_FRUSTUM:
                bvh_leaf<SEG_FRUSTUM, REFINE, COUNT>(L, q[k], c[k], cnt, R, tb, ib, bp, pd, slot, key0, fl, ev);
                break;
            case SEG_SPHERE:
                bvh_leaf<SEG_SPHERE, REFINE, COUNT>(L, q[k], c[k], cnt, R, tb, ib, bp, pd, slot, key0, fl, ev);
                break;
            case SEG_RING:
                bvh_leaf<SEG_RING, REFINE, COUNT>(L, q[k], c[k], cnt, R, tb, ib, bp, pd, slot, key0, fl, ev);
                break;
            default:
                bvh_leaf<SEG_RECT, REFINE, COUNT>(L, q[k], c[k], cnt, R, tb, ib, bp, pd, slot, key0, fl, ev);
                break;
            }
        }
    }
}

// One ray slot's walk of the hierarchy from `root` (-1: none).
template <bool REFINE, bool COUNT>
__device__ __forceinline__ void bvh_walk(GPTR(BvhNode) N, GPTR(FastPrim) L, int root, int key0, const Ray32 &R,
                                         const BoxRay &B, float &tb, int &ib, Pending &pd, int slot, uint32_t *stk,
                                         uint32_t &fl, uint32_t &ev)
{
    int bp = -1, node = root, sp = 0;
    BvhNode nd;
    if (node >= 0) nd = ldg(N + node);
    while (node >= 0) bvh_step<REFINE, COUNT>(N, L, node, nd, sp, bp, key0, R, B, tb, ib, pd, slot, stk, fl, ev);
}

// FP64 form, a scene whose hierarchy holds every primitive (bvh_on == 2: no
// band stacks): the walk with the FP64 test of each leaf's record in the
// caller's order (hit_d, as find_d), the nearest hit and, at equal distance,
// the lowest index: find_d's result (strict < in the caller's order).  The
// boxes are crossed in FP32 on the rounded ray, whose rounding moves it by
// far less than the boxes' padding, against t_best rounded to FP32 (the same
// margin argument as the fast form's walk).
__device__ __forceinline__ int find_d_bvh(CPTR(GeomDesc) G, const Ray64 &ray, uint32_t *stk, uint64_t &sflops,
                                          uint64_t &ev)
{
    GPTR(BvhNode) N = (GPTR(BvhNode))G->bvh_nodes;
    GPTR(FastPrim) L = (GPTR(FastPrim))G->bvh_prims;
    GPTR(rf_primitive) P = (GPTR(rf_primitive))G->prims;
    Ray32 r32;
    r32.sx = (float)ray.sx; r32.sy = (float)ray.sy; r32.sz = (float)ray.sz;
    r32.cx = (float)ray.cx; r32.cy = (float)ray.cy; r32.cz = (float)ray.cz;
    const BoxRay B = box_ray(r32, G);
    double tbest = dinf();
    int obj = -1;
    uint32_t tests = 0, fl = kFlopsBoxRay;
    int node = G->bvh_root, sp = 0;
    while (node >= 0) {
        const BvhNode nd = ldg(N + node);
        const float tb = (float)tbest;
        float e0, l0, e1, l1;
        box_span(B, nd, 0, tb, e0, l0);
        box_span(B, nd, 1, tb, e1, l1);
        const int ci[2] = {bvh_index(nd.ref[0]), bvh_index(nd.ref[1])};
        const int ni[2] = {bvh_leafword(nd.ref[0]), bvh_leafword(nd.ref[1])};
        const bool h0 = e0 <= l0;
        const bool h1 = (e1 <= l1) & (ci[1] >= 0);
        tests += 2u;
        fl += 2u * kFlopsBox;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if ((k ? h1 : h0) && ni[k] > 0) {
                const int c = ci[k], cnt = ni[k] & 7;
                for (int j = c; j < c + cnt; ++j) {
                    // the record's caller index (FastPrim::orig)
                    const int w = ((GPTR(int32_t))(L + j))[offsetof(FastPrim, orig) / 4];
                    const rf_primitive p = ldg(P + w);
                    const double t = hit_d(p.type, p.invm, p.util1, p.util2, ray);
                    if (t < tbest || (t == tbest && w < obj)) { tbest = t; obj = w; }
                    ++tests;
                    fl += (uint32_t)flops_per_test(p.type);
                }
            }
        }
        const bool i0 = h0 & (ni[0] == 0), i1 = h1 & (ni[1] == 0);
        if (i0 & i1) {
            const bool near0 = e0 <= e1;
            stk[sp * blockDim.x] = (uint32_t)(near0 ? ci[1] : ci[0]);
            ++sp;
            node = near0 ? ci[0] : ci[1];
        } else if (i0) {
            node = ci[0];
        } else if (i1) {
            node = ci[1];
        } else {
            node = sp > 0 ? (int)stk[--sp * blockDim.x] : -1;
        }
    }
    ev += tests;
    sflops += fl;
    return obj;
}

// The hierarchy form of scan_fast: band stacks from the run list as there,
// then per ray slot (those set in vmask) the hierarchy.
// stk: the lane's traversal stack in LDS (the deepest hierarchy's depth + 1
// entries, blockDim.x apart, sized by the host).
template <int NR, bool REFINE, bool COUNT>
__device__ __forceinline__ void scan_bvh(CPTR(GeomDesc) G, const Ray32 (&R)[NR], float (&tb)[NR], int (&ib)[NR],
                                         Pending &pd, int stacks, uint32_t &fl, uint32_t &ev, uint32_t vmask,
                                         uint32_t *stk)
{
    CPTR(FastPrim) F = cst(G->fast);
    const int n_runs = G->n_cyl_runs;
    for (int j = 0; j < n_runs; ++j) {
        const CylRun run = ldc(cst(G->cyl_runs) + j);
        if (run.stack)
            stack_loop<NR, REFINE, COUNT, false>(F, cst(G->bands), run, stacks >= 2 ? run.stack : 1, R, tb, ib, pd,
                                                 fl, ev, vmask);
    }
    GPTR(BvhNode) N = (GPTR(BvhNode))G->bvh_nodes;
    GPTR(FastPrim) L = (GPTR(FastPrim))G->bvh_prims;
    const int root = G->bvh_root, key0 = G->seg_begin[SEG_CYL] * 8;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const bool act = (vmask >> i) & 1u;
        const BoxRay B = box_ray(R[i], G);
        if (COUNT && act) fl += kFlopsBoxRay;
        bvh_walk<REFINE, COUNT>(N, L, act ? root : -1, key0, R[i], B, tb[i], ib[i], pd, i, stk, fl, ev);
    }
}

// stacks: the cylinder segment runs as its CylRun list (band stacks share the
// quadric; 2: sorted stacks are searched); otherwise as one per-test run.
// fl and ev gain the band stacks' flops and evaluations of the ray slots set
// in vmask (the per-test primitives are counted on the host).
template <int NR, bool REFINE, bool COUNT = true, bool BT = false>
__device__ __forceinline__ void scan_fast(CPTR(GeomDesc) G, const Ray32 (&R)[NR], float (&tb)[NR], int (&ib)[NR],
                                          Pending &pd, int stacks, uint32_t &fl, uint32_t &ev, uint32_t vmask,
                                          BandsLds BL = BandsLds{})
{
    CPTR(FastPrim) F = cst(G->fast);
    const int n_runs = stacks ? G->n_cyl_runs : 1;
    for (int j = 0; j < n_runs; ++j) {
        CylRun run{G->seg_begin[SEG_CYL], G->seg_end[SEG_CYL], 0, 0};
        if (stacks) run = ldc(cst(G->cyl_runs) + j);
        if (run.stack)
            stack_loop<NR, REFINE, COUNT, BT>(F, cst(G->bands), run, stacks >= 2 ? run.stack : 1, R, tb, ib, pd, fl,
                                              ev, vmask, BL);
        else
            seg_loop<SEG_CYL, NR, REFINE, false>(F, run.begin, run.end, R, tb, ib, pd);
    }
    seg_loop<SEG_FRUSTUM, NR, REFINE, false>(F, G->seg_begin[SEG_FRUSTUM], G->seg_end[SEG_FRUSTUM], R, tb, ib, pd);
    seg_loop<SEG_SPHERE, NR, REFINE, false>(F, G->seg_begin[SEG_SPHERE], G->seg_end[SEG_SPHERE], R, tb, ib, pd);
    zring_loop<NR>(F, G->seg_begin[SEG_RING], G->zring_end, R, tb, ib, pd.eps);
    seg_loop<SEG_RING, NR, REFINE, false>(F, G->zring_end, G->seg_end[SEG_RING], R, tb, ib, pd);
    seg_loop<SEG_RECT, NR, REFINE, false>(F, G->seg_begin[SEG_RECT], G->seg_end[SEG_RECT], R, tb, ib, pd);
}

// Rays 0, 1 (upper = false) or 2, 3 of a work item whose four uniforms are
// f: the upper half takes the uniforms rotated by two more places.  Four
// selects on one condition, materialised, so that the compiler does not turn
// them into a per-lane index select for each of the eight uses.
__device__ __forceinline__ void emit_half(const float (&f)[4], bool upper, const Emitter &E, Ray32 (&R)[2])
{
    float g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        g[k] = upper ? f[(k + 2) & 3] : f[k];
        asm volatile("" : "+v"(g[k]));
    }
    float r4[4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) r4[j] = g[(i + j + 1) & 3];
        R[i] = emit_fast(&E, r4);
    }
}

// The NR rays of ray group `base` (unit-local ray index) of one emitter.
// base is a multiple of NR (NR = 1, 2, 4, 8), so a group never straddles two
// work items unless NR > 4, and each ray's rotation of its work item's four
// uniforms (ray v of a work item takes f[v+1], f[v+2], f[v+3], f[v], indices
// mod 4) is known at compile time up to, for NR = 2, which half of the work
// item the group is: one Philox block per work item and no per-ray
// selection of the rotation.
template <int NR>
__device__ __forceinline__ void emit_group(const KParams &kp, uint32_t gid, int emitter, int64_t wi0, int64_t base,
                                           const Emitter &E, Ray32 (&R)[NR])
{
    static_assert(NR == 1 || NR == 2 || NR % 4 == 0, "rays per thread");
    float f[4];
    if constexpr (NR == 1) {
        workitem_floats(kp, gid, emitter, wi0 + (base >> 2), f);
        float r4[4];
        rotate_rands(f, (int)(base & 3), r4);
        R[0] = emit_fast(&E, r4);
    } else if constexpr (NR == 2) {
        workitem_floats(kp, gid, emitter, wi0 + (base >> 2), f);
        // rays v = 0, 1 of the work item, or v = 2, 3: the upper half takes the
        // uniforms rotated by two more places.  Four selects on one condition,
        // materialised, so that the compiler does not turn them into a
        // per-lane index select for each of the eight uses.
        const bool upper = (base & 2) != 0;
        float g[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            g[k] = upper ? f[(k + 2) & 3] : f[k];
            asm volatile("" : "+v"(g[k]));
        }
        float r4[4];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) r4[j] = g[(i + j + 1) & 3];
            R[i] = emit_fast(&E, r4);
        }
    } else {
#pragma unroll
        for (int q = 0; q < NR / 4; ++q) {
            workitem_floats(kp, gid, emitter, wi0 + (base >> 2) + q, f);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float r4[4];
                rotate_rands(f, v, r4);
                R[4 * q + v] = emit_fast(&E, r4);
            }
        }
    }
}

// Child scan of the fast BV kernel over the caller's order [start, end).
template <bool REFINE>
__device__ __forceinline__ int find_fast_range(CPTR(FastPrim) F, CPTR(rf_primitive) P, int start, int end,
                                               const Ray32 &R, uint32_t &nref, float tau)
{
    int obj = -1;
    float tbest = finf();
    for (int w = start; w < end; ++w) {
        const FastPrim q = ldc(F + w);
        bool f;
        float t = fast_test_any<REFINE>(q, R, f, tau, EPS_F);
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
__device__ __forceinline__ int find_bv_fast(CPTR(GeomDesc) G, const Ray32 &R, uint64_t &tests, uint32_t &nref,
                                            float tau)
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
        float t = fast_test_any<REFINE>(q, R, f, tau, EPS_F);
        if constexpr (REFINE) {
            if (__builtin_amdgcn_ballot_w64(f)) {
                if (f) { t = refine_bv(B, w, R); ++nref; }
            }
        }
        ++tests;
        if (t < tbest) {
            if (meta.settings == 1) {
                objectIndex = find_fast_range<REFINE>(F, P, meta.start, meta.end, R, nref, tau);
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
            objectIndex = find_fast_range<REFINE>(F, P, meta.start, meta.end, R, nref, tau);
            tests += (uint64_t)(meta.end - meta.start);
            pending = false;
        }
    }
    return objectIndex;
}

// ---------------------------------------------------------------------------
// Work units and the tally
// ---------------------------------------------------------------------------
// The hit matrix a launch tallies into: the caller's (kp.hits) when set, else
// the geometry's own.
__device__ __forceinline__ unsigned long long *hits_of(const KParams &kp, CPTR(GeomDesc) G)
{
    return kp.hits ? kp.hits : G->hits;
}

struct UnitRef {
    CPTR(GeomDesc) G;
    int emitter;
    int64_t wi0, n_wi;
    uint32_t gid;
};

// Unit u of a persistent launch: from the unit array, or implicit -- work
// items [imp_wi0 + c imp_unit, ...) of emitter imp_emitter, or, when
// imp_emitter < 0, of emitter u / imp_upe (chunk c = u mod imp_upe; every
// emitter of geoms[0] has the same budget imp_n_wi).
__device__ __forceinline__ UnitRef unit_of(const KParams &kp, int u)
{
    UnitRef r;
    if (kp.units) {
        const WorkUnit U = ldc(cst(kp.units) + u);
        r.G = cst(kp.geoms) + U.geom;
        r.emitter = U.emitter;
        r.wi0 = U.wi0;
        r.n_wi = U.n_wi;
        r.gid = r.G->geometry_id;
        return r;
    }
    int c = u;
    r.emitter = kp.imp_emitter;
    if (kp.imp_emitter < 0) {
        r.emitter = u / kp.imp_upe;
        c = u - r.emitter * kp.imp_upe;
    }
    r.G = cst(kp.geoms);
    r.wi0 = kp.imp_wi0 + (int64_t)c * kp.imp_unit;
    const int64_t left = kp.imp_wi0 + kp.imp_n_wi - r.wi0;
    r.n_wi = left < kp.imp_unit ? left : kp.imp_unit;
    r.gid = kp.imp_geom_id;
    return r;
}

// Flush of a unit's LDS bins into its emitter's row: one 64-bit atomic add
// per non-zero bin (a row may have several units; plain stores of whole rows
// measured no less traffic and no faster).
__device__ __forceinline__ void flush_row(const KParams &kp, const UnitRef &ur, const uint32_t *bins)
{
    const int n = ur.G->n;
    unsigned long long *row = hits_of(kp, ur.G) + (int64_t)ur.emitter * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t c = bins[i];
        if (c) atomicAdd(&row[i], (unsigned long long)c);
    }
}
template <bool TRACE, bool DIRECT>
__device__ __forceinline__ void tally(const KParams &kp, CPTR(GeomDesc) G, int emitter, int64_t ray_abs,
                                      int idx, uint32_t *bins)
{
    if constexpr (TRACE) {
        kp.trace_out[ray_abs] = idx;
    } else {
        if (idx < 0) return;
        if (DIRECT || kp.bins == 0)
            atomicAdd(&hits_of(kp, G)[(int64_t)emitter * G->n + idx], 1ull);
        else
            atomicAdd(&bins[idx], 1u);
    }
}

// Direction cell of a ray for the hierarchy form's tile sort: the octant
// (sign bits) and, on the octahedron |x| + |y| + |z| = 1, |x| and |y| in
// eighths, interleaved (Morton order, so that consecutive cells are
// neighbours); 64 cells per octant.  Cell BVH_CELLS - 1 is never taken (|x|
// and |y| cannot both reach seven eighths): it sorts the tile's missing rays
// last.
constexpr int BVH_CELLS = 512;
constexpr int BVH_TILE_PASSES = 4;   // sub-passes of blockDim.x x NR rays per sorted tile

__device__ __forceinline__ int dir_cell(const Ray32 &R)
{
    const float ax = __builtin_fabsf(R.cx), ay = __builtin_fabsf(R.cy), az = __builtin_fabsf(R.cz);
    const float q = 8.0f * rcp_fast(ax + ay + az);
    const int qa = min(7, (int)(ax * q)), qb = min(7, (int)(ay * q));
    const int oct = (R.cx < 0.0f ? 1 : 0) | (R.cy < 0.0f ? 2 : 0) | (R.cz < 0.0f ? 4 : 0);
    const int m = (qa & 1) | ((qb & 1) << 1) | ((qa & 2) << 1) | ((qb & 2) << 2) | ((qa & 4) << 2) | ((qb & 4) << 3);
    return oct * 64 + m;
}

// Unit body of the hierarchy form (MATH: fast, or FP64 on a scene whose
// hierarchy holds every primitive).  The group takes its rays in tiles of
// BVH_TILE_PASSES x blockDim.x x NR, counting-sorts each tile by direction
// cell in LDS, and walks it in BVH_TILE_PASSES sub-passes in which lane t
// holds sorted rays t, t + blockDim.x, ... of the sub-pass: the lanes of a
// wave walk rays of one emitter in neighbouring directions, which visit much
// the same nodes.  A ray is emitted again from its index after the sort (the
// same ray, as for an FP64 re-solve) and tallied under its own index; the
// order within a cell does not matter, as no ray's result depends on its lane.
template <int MATH, bool TRACE, int NR, bool REFINE, bool COUNT>
__device__ __forceinline__ void run_unit_bvh(const KParams &kp, CPTR(GeomDesc) G, const Emitter &E, uint32_t gid,
                                             int emitter, int64_t wi0, int64_t n_wi, uint32_t *bins, uint32_t *pend,
                                             uint64_t &sflops, uint64_t &sevals, uint32_t &nref, uint32_t &nover,
                                             uint32_t *bstk)
{
    static_assert(NR <= 8, "tile of at most 8 x 256 rays per sub-pass");
    constexpr int TMAX = BVH_TILE_PASSES * 256 * NR;
    __shared__ uint32_t s_cnt[BVH_CELLS];
    __shared__ uint16_t s_ray[TMAX];
    __shared__ uint16_t s_cell[TMAX];
    const int64_t ray0 = (wi0 - kp.trace_wi0) * 4;
    const int64_t n_rays = n_wi * 4;
    const int sub = (int)blockDim.x * NR;
    const int tile = BVH_TILE_PASSES * sub;
    uint32_t fl = 0, ev = 0;
    for (int64_t p0 = 0; p0 < n_rays; p0 += tile) {
        for (int k = threadIdx.x; k < BVH_CELLS; k += blockDim.x) s_cnt[k] = 0u;
        __syncthreads();
        // cells of this lane's rays (NR consecutive per sub-pass) and the histogram
        for (int sp = 0; sp < BVH_TILE_PASSES; ++sp) {
            const int own = sp * sub + (int)threadIdx.x * NR;
            Ray32 R[NR];
            if constexpr (MATH == KM_FAST) {
                emit_group<NR>(kp, gid, emitter, wi0, p0 + own, E, R);
            } else {
#pragma unroll
                for (int i = 0; i < NR; ++i) {
                    const int64_t r = p0 + own + i;
                    float f[4], r4[4];
                    workitem_floats(kp, gid, emitter, wi0 + (r >> 2), f);
                    rotate_rands(f, (int)(r & 3), r4);
                    const double d4[4] = {r4[0], r4[1], r4[2], r4[3]};
                    const Ray64 rd = emit_d(&E, d4);
                    R[i].cx = (float)rd.cx;
                    R[i].cy = (float)rd.cy;
                    R[i].cz = (float)rd.cz;
                }
            }
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                const int c = p0 + own + i < n_rays ? dir_cell(R[i]) : BVH_CELLS - 1;
                s_cell[own + i] = (uint16_t)c;
                atomicAdd(&s_cnt[c], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            // exclusive prefix over the cells, BVH_CELLS / 64 per lane of the first wave
            constexpr int PER = BVH_CELLS / 64;
            uint32_t c[PER], sum = 0;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                c[k] = s_cnt[PER * threadIdx.x + k];
                sum += c[k];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(incl, d);
                if ((int)threadIdx.x >= d) incl += o;
            }
            uint32_t run = incl - sum;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                s_cnt[PER * threadIdx.x + k] = run;
                run += c[k];
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < tile; k += blockDim.x) s_ray[atomicAdd(&s_cnt[s_cell[k]], 1u)] = (uint16_t)k;
        __syncthreads();
        for (int sp = 0; sp < BVH_TILE_PASSES; ++sp) {
            if constexpr (MATH == KM_FP64) {
                // FP64: each sorted ray emitted again, walked and tallied
#pragma unroll
                for (int i = 0; i < NR; ++i) {
                    const int64_t r = p0 + s_ray[sp * sub + (int)threadIdx.x + i * (int)blockDim.x];
                    if (r >= n_rays) continue;
                    float f[4], r4[4];
                    workitem_floats(kp, gid, emitter, wi0 + (r >> 2), f);
                    rotate_rands(f, (int)(r & 3), r4);
                    const double d4[4] = {r4[0], r4[1], r4[2], r4[3]};
                    const Ray64 rd = emit_d(&E, d4);
                    const int idx = find_d_bvh(G, rd, bstk, sflops, sevals);
                    tally<TRACE, false>(kp, G, emitter, ray0 + r, idx, bins);
                }
                continue;
            }
            Ray32 R[NR];
            int idx[NR];
            float tb[NR];
            int ib[NR];
            uint32_t vmask = 0;
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                idx[i] = s_ray[sp * sub + (int)threadIdx.x + i * (int)blockDim.x];
                const int64_t r = p0 + idx[i];
                if (r < n_rays) vmask |= 1u << i;
                float f[4], r4[4];
                workitem_floats(kp, gid, emitter, wi0 + (r >> 2), f);
                rotate_rands(f, (int)(r & 3), r4);
                R[i] = emit_fast(&E, r4);
                tb[i] = finf();
                ib[i] = -1;
            }
            Pending pd{pend, 0, NO_KEY, 0, vcopy(kp.tau), vcopy(EPS_F)};
            scan_bvh<NR, REFINE, COUNT>(G, R, tb, ib, pd, kp.stacks, fl, ev, vmask, bstk);
            if (COUNT && ((fl | ev) >> 30)) {
                sflops += fl;
                sevals += ev;
                fl = ev = 0;
            }
            if constexpr (REFINE) {
                const auto ray_of = [&](int k) {
                    int v = idx[0];
#pragma unroll
                    for (int i = 1; i < NR; ++i) v = k == i ? idx[i] : v;
                    return p0 + v;
                };
                if (__builtin_amdgcn_ballot_w64(pd.count > 0))
                    resolve_pending<NR>(kp, G, E, gid, emitter, wi0, ray_of, tb, ib, pd, nref);
                if (__builtin_amdgcn_ballot_w64(pd.resume != NO_KEY))
                    finish_dropped<NR>(kp, G, E, gid, emitter, wi0, ray_of, cst(G->fast), G->seg_begin, G->seg_end,
                                       tb, ib, pd, nref, nover);
            }
#pragma unroll
            for (int i = 0; i < NR; ++i)
                if ((vmask >> i) & 1u) tally<TRACE, false>(kp, G, emitter, ray0 + p0 + idx[i], ib[i], bins);
        }
    }
    if constexpr (COUNT) {
        sflops += fl;
        sevals += ev;
    }
}

// ---------------------------------------------------------------------------
// Unit body: the rays of work items [lo, hi) (unit-local) handled by this
// thread, stepping by `stride` work items.
// ---------------------------------------------------------------------------
template <int MATH, bool BV, bool TRACE, int NR, bool REFINE, bool DIRECT, bool COUNT = true, bool BT = false,
          bool BVH = false>
__device__ __forceinline__ void run_unit(const KParams &kp, CPTR(GeomDesc) G, uint32_t gid, int emitter,
                                         int64_t wi0, int64_t n_wi, int64_t first, int64_t stride,
                                         uint32_t *bins, uint32_t *pend, uint64_t &tests, uint64_t &sflops,
                                         uint64_t &sevals, uint32_t &nref, uint32_t &nover, BandsLds BL = BandsLds{},
                                         uint32_t *bstk = nullptr)
{
    const Emitter E = ldc(cst(G->emitters) + emitter);
    const int64_t ray0 = (wi0 - kp.trace_wi0) * 4;   // absolute ray index base (trace launches)
    if constexpr (MATH == KM_FP64 && !BV && !DIRECT && BVH) {
        // FP64, a scene whose hierarchy holds every primitive: sorted tiles
        // of 2 rays per thread
        if (G->bvh_on == 2) {
            run_unit_bvh<KM_FP64, TRACE, 2, false, true>(kp, G, E, gid, emitter, wi0, n_wi, bins, pend, sflops, sevals,
                                                         nref, nover, bstk);
            return;
        }
    }
    if constexpr (MATH == KM_FAST && !BV && !DIRECT) {
        // a scene with a hierarchy, in a launch of the hierarchy form
        if (BVH && G->bvh_on) {
            run_unit_bvh<KM_FAST, TRACE, NR, REFINE, COUNT>(kp, G, E, gid, emitter, wi0, n_wi, bins, pend, sflops,
                                                            sevals, nref, nover, bstk);
            return;
        }
        // NR rays per thread: ray groups first, first+stride, ... of NR
        // consecutive rays; with 2 rays per thread, work items first,
        // first+stride, ... whose two halves are consecutive groups of the
        // thread (one Philox block for both)
        const int64_t n_rays = n_wi * 4;
        // the unit's band-stack flops and evaluations, in 32 bits (folded into
        // the 64-bit totals before they could wrap)
        uint32_t fl = 0, ev = 0;
        // (kept in LDS between the halves, not in registers across the scan)
        constexpr bool PAIR = NR == 2;
        __shared__ float s_fw[PAIR ? 4 * 256 : 1];
        const int64_t step = PAIR ? stride * 4 : stride * NR;
        int half = 1;
        for (int64_t base = PAIR ? first * 4 : first * NR; base < n_rays; base += PAIR ? (half ? step - 2 : 2) : step) {
            Ray32 R[NR];
            float tb[NR];
            int ib[NR];
            if constexpr (PAIR) {
                half ^= 1;
                float fw[4];
                if (!half) {
                    workitem_floats(kp, gid, emitter, wi0 + (base >> 2), fw);
#pragma unroll
                    for (int k = 0; k < 4; ++k) s_fw[k * blockDim.x + threadIdx.x] = fw[k];
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k) fw[k] = s_fw[k * blockDim.x + threadIdx.x];
                }
                emit_half(fw, half != 0, E, R);
            } else {
                emit_group<NR>(kp, gid, emitter, wi0, base, E, R);
            }
#pragma unroll
            for (int i = 0; i < NR; ++i) { tb[i] = finf(); ib[i] = -1; }
            Pending pd{pend, 0, NO_KEY, 0, vcopy(kp.tau), vcopy(EPS_F)};
            const int64_t left = n_rays - base;
            const uint32_t vmask = left < NR ? (1u << left) - 1u : (1u << NR) - 1u;
            scan_fast<NR, REFINE, COUNT, BT>(G, R, tb, ib, pd, kp.stacks, fl, ev, vmask, BL);
            if (COUNT && ((fl | ev) >> 30)) {
                sflops += fl;
                sevals += ev;
                fl = ev = 0;
            }
            if constexpr (REFINE) {
                const auto ray_of = [base](int k) { return base + k; };
                if (__builtin_amdgcn_ballot_w64(pd.count > 0))
                    resolve_pending<NR>(kp, G, E, gid, emitter, wi0, ray_of, tb, ib, pd, nref);
                if (__builtin_amdgcn_ballot_w64(pd.resume != NO_KEY))
                    finish_dropped<NR>(kp, G, E, gid, emitter, wi0, ray_of, cst(G->fast), G->seg_begin, G->seg_end,
                                       tb, ib, pd, nref, nover);
            }
#pragma unroll
            for (int i = 0; i < NR; ++i)
                if (base + i < n_rays) tally<TRACE, DIRECT>(kp, G, emitter, ray0 + base + i, ib[i], bins);
        }
        if constexpr (COUNT) {
            sflops += fl;
            sevals += ev;
        }
    } else {
        // one work item (4 rays) per thread and step
        for (int64_t wl = first; wl < n_wi; wl += stride) {
            float f[4];
            workitem_floats(kp, gid, emitter, wi0 + wl, f);
            if constexpr (MATH == KM_COMPAT && !BV) {
                // the work item's 4 rays share each primitive visit
                Ray32 R4[4];
                int idx4[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float r4[4];
                    rotate_rands(f, v, r4);
                    R4[v] = emit_c(&E, r4);
                }
                uint32_t ev = 0;
                sflops += find_c_multi<4>(G, R4, idx4, kp.stacks, ev);
                sevals += ev;
#pragma unroll
                for (int v = 0; v < 4; ++v) tally<TRACE, DIRECT>(kp, G, emitter, ray0 + wl * 4 + v, idx4[v], bins);
                continue;
            }
            for (int v = 0; v < 4; ++v) {
                float r4[4];
                rotate_rands(f, v, r4);
                int idx;
                if constexpr (MATH == KM_COMPAT) {
                    const Ray32 ray = emit_c(&E, r4);
                    idx = BV ? find_bv_cd<KM_COMPAT>(G, ray, tests, kp.stacks)
                             : find_c(cst(G->prims), 0, G->n, ray);
                } else if constexpr (MATH == KM_FP64) {
                    const double d4[4] = {r4[0], r4[1], r4[2], r4[3]};
                    const Ray64 ray = emit_d(&E, d4);
                    idx = BV ? find_bv_cd<KM_FP64>(G, ray, tests, kp.stacks)
                             : find_d_runs(G, 0, G->n, ray, kp.stacks, sflops, sevals);
                } else {
                    const Ray32 ray = emit_fast(&E, r4);
                    if constexpr (BV) {
                        idx = find_bv_fast<REFINE>(G, ray, tests, nref, kp.tau);
                    } else {
                        Ray32 R1[1] = {ray};
                        float tb1[1] = {finf()};
                        int ib1[1] = {-1};
                        Pending pd{pend, 0, NO_KEY, 0, vcopy(kp.tau), vcopy(EPS_F)};
                        uint32_t fl1 = 0, ev1 = 0;
                        scan_fast<1, REFINE>(G, R1, tb1, ib1, pd, kp.stacks, fl1, ev1, 1u);
                        sflops += fl1;
                        sevals += ev1;
                        if constexpr (REFINE) {
                            const int64_t r0 = wl * 4 + v;
                            const auto ray_of = [r0](int k) { return r0 + k; };
                            if (__builtin_amdgcn_ballot_w64(pd.count > 0))
                                resolve_pending<1>(kp, G, E, gid, emitter, wi0, ray_of, tb1, ib1, pd, nref);
                            if (__builtin_amdgcn_ballot_w64(pd.resume != NO_KEY))
                                finish_dropped<1>(kp, G, E, gid, emitter, wi0, ray_of, cst(G->fast), G->seg_begin,
                                                  G->seg_end, tb1, ib1, pd, nref, nover);
                        }
                        idx = ib1[0];
                    }
                }
                tally<TRACE, DIRECT>(kp, G, emitter, ray0 + wl * 4 + v, idx, bins);
            }
        }
    }
}

// stats[0]: tests counted on the device (BV forms: surface tests and child
// evaluations); stats[1]: pairs re-solved in FP64; stats[2]: lane re-scans
// after a full re-solve list; stats[3]: algorithmic flops counted on the
// device (band stacks of the plain scan; all child scans of the binned BV
// kernel); stats[4]: band-stack evaluations of the plain scans (a quadric per
// ray and stack, a z-range check per band looked at).
__device__ __forceinline__ void flush_stats(const KParams &kp, uint64_t tests, uint32_t nref, uint32_t nover,
                                            uint64_t flops = 0, uint64_t evals = 0)
{
    __shared__ unsigned long long s_stats[kNStats];
    if (threadIdx.x < kNStats) s_stats[threadIdx.x] = 0;
    __syncthreads();
    if (tests) atomicAdd(&s_stats[0], (unsigned long long)tests);
    if (nref) atomicAdd(&s_stats[1], (unsigned long long)nref);
    if (nover) atomicAdd(&s_stats[2], (unsigned long long)nover);
    if (flops) atomicAdd(&s_stats[3], (unsigned long long)flops);
    if (evals) atomicAdd(&s_stats[4], (unsigned long long)evals);
    __syncthreads();
    if (threadIdx.x < kNStats && kp.stats && s_stats[threadIdx.x])
        atomicAdd(&kp.stats[threadIdx.x], s_stats[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// fast plain form with the ray transform on the matrix cores
//
// A wave traces 64 rays per pass as 4 sets of 16.  Primitives come in quads;
// one coalesced 256-byte load per quad gives every lane one element of the
// 16x4 A operand (4 primitives x 4 output rows x [x, y, z, 1]).  For each
// set, two v_mfma_f32_16x16x4f32 transform the set's origins (B column
// [s; 1]) and directions ([c; 0]); lane l then holds the four object-space
// rows of ray (set, l % 16) for primitive slot l / 16 and runs only the
// quadric / plane arithmetic on the VALU.  Each lane keeps the nearest hit
// of its slot; a cross-lane minimum over the 4 slots finishes each ray.
// Output rows per type (quads built on the host):
//   cylinder  x, y, z-1/2, 0          frustum  x, y, w = 1+(u1-1)z, z-1/2
//   sphere    x, y, z, 0              disc / annulus  x/ro, y/ro, z, (ri/ro)^2
//   rectangle x, y, z, 0              padding  never hits or flags
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int KIND>
__device__ __forceinline__ float quad_test(const f32x4 &o, const f32x4 &d, bool &flag, float tau, float eps)
{
    if constexpr (KIND == SEG_CYL) {
        const float sx = o.x, sy = o.y, sz = o.z, cx = d.x, cy = d.y, cz = d.z;
        const float A = __builtin_fmaf(cx, cx, cy * cy);
        const float b = __builtin_fmaf(sx, cx, sy * cy);
        const float w = __builtin_fmaf(sx, cy, -(sy * cx));
        const float D = __builtin_fmaf(-w, w, A);
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float tm = -b * rA;
        const float t1 = __builtin_fmaf(-sq, rA, tm);
        const float t2 = __builtin_fmaf(sq, rA, tm);
        const float z1 = __builtin_fmaf(cz, t1, sz);
        const float z2 = __builtin_fmaf(cz, t2, sz);
        const bool e1 = t1 > eps, e2 = t2 > eps;
        const bool v1 = e1 & (__builtin_fabsf(z1) <= 0.5f);
        const bool v2 = e2 & (__builtin_fabsf(z2) <= 0.5f);
        const float zm = __builtin_fmaf(cz, tm, sz);
        flag = ((__builtin_fabsf(D) <= tau * A) & (__builtin_fabsf(zm) <= 0.5f + tau)) |
               (e1 & (__builtin_fabsf(__builtin_fabsf(z1) - 0.5f) <= tau)) |
               (e2 & (__builtin_fabsf(__builtin_fabsf(z2) - 0.5f) <= tau));
        return v1 ? t1 : (v2 ? t2 : finf());
    } else if constexpr (KIND == SEG_FRUSTUM) {
        const float sx = o.x, sy = o.y, sw = o.z, sz = o.w, cx = d.x, cy = d.y, cw = d.z, cz = d.w;
        const float Axy = __builtin_fmaf(cx, cx, cy * cy);
        const float A = __builtin_fmaf(-cw, cw, Axy);
        const float b = __builtin_fmaf(sx, cx, __builtin_fmaf(sy, cy, -(sw * cw)));
        const float C = __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, -(sw * sw)));
        const float D = __builtin_fmaf(b, b, -(A * C));
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float tm = -b * rA;
        const float ta = __builtin_fmaf(-sq, rA, tm);
        const float tb = __builtin_fmaf(sq, rA, tm);
        const float t1 = __builtin_fminf(ta, tb);
        const float t2 = __builtin_fmaxf(ta, tb);
        const float z1 = __builtin_fmaf(cz, t1, sz);
        const float z2 = __builtin_fmaf(cz, t2, sz);
        const bool e1 = t1 > eps, e2 = t2 > eps;
        const bool v1 = e1 & (__builtin_fabsf(z1) <= 0.5f);
        const bool v2 = e2 & (__builtin_fabsf(z2) <= 0.5f);
        const float zm = __builtin_fmaf(cz, tm, sz);
        flag = ((__builtin_fabsf(D) <= tau * (b * b)) & (__builtin_fabsf(zm) <= 0.5f + tau)) |
               (__builtin_fabsf(A) < tau * Axy) |
               (e1 & (__builtin_fabsf(__builtin_fabsf(z1) - 0.5f) <= tau)) |
               (e2 & (__builtin_fabsf(__builtin_fabsf(z2) - 0.5f) <= tau));
        return v1 ? t1 : (v2 ? t2 : finf());
    } else if constexpr (KIND == SEG_SPHERE) {
        const float sx = o.x, sy = o.y, sz = o.z, cx = d.x, cy = d.y, cz = d.z;
        const float A = __builtin_fmaf(cx, cx, __builtin_fmaf(cy, cy, cz * cz));
        const float b = __builtin_fmaf(sx, cx, __builtin_fmaf(sy, cy, sz * cz));
        const float C = __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, __builtin_fmaf(sz, sz, -1.0f)));
        const float D = __builtin_fmaf(b, b, -(A * C));
        const float sq = sqrt_fast(D);
        const float rA = rcp_fast(A);
        const float t1 = (-b - sq) * rA;
        const float t2 = (sq - b) * rA;
        flag = (__builtin_fabsf(D) <= tau * (b * b)) & (b < 0.0f);
        return (t1 > eps) ? t1 : ((t2 > eps) ? t2 : finf());
    } else {
        const float t = -o.z * rcp_fast(d.z);
        const float hx = __builtin_fmaf(d.x, t, o.x);
        const float hy = __builtin_fmaf(d.y, t, o.y);
        flag = false;
        bool v;
        if constexpr (KIND == SEG_RING) {
            const float rr = __builtin_fmaf(hx, hx, hy * hy);
            v = (t > eps) & (rr <= 1.0f) & (rr >= o.w);
        } else {
            v = (t > eps) & (__builtin_fabsf(hx) <= 1.0f) & (__builtin_fabsf(hy) <= 1.0f);
        }
        return v ? t : finf();
    }
}

template <int KIND, bool REFINE, bool RESCAN>
__device__ __forceinline__ void quad_loop(const float *__restrict__ Q, int qb, int qe, const float (&Bo)[4],
                                          const float (&Bd)[4], float (&tb)[4], int (&ib)[4], Pending &pd,
                                          int lane)
{
    if (qb >= qe) return;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    float a_next = Q[(size_t)qb * 64 + lane];
    for (int q = qb; q < qe; ++q) {
        const float a = a_next;
        if (q + 1 < qe) a_next = Q[(size_t)(q + 1) * 64 + lane];
        const int pos = 4 * q + (lane >> 4);
        f32x4 Do[4], Dd[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            Do[st] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bo[st], zero, 0, 0, 0);
            Dd[st] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bd[st], zero, 0, 0, 0);
        }
        float t[4];
        bool f[4];
        bool any = false;
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            t[st] = quad_test<KIND>(Do[st], Dd[st], f[st], pd.tau, pd.eps);
            if constexpr (RESCAN) {
                const bool skip = q * 4 + st < pd.from;
                f[st] = f[st] & !skip;
                t[st] = skip ? finf() : t[st];
            }
            any |= f[st];
        }
        if constexpr (REFINE && KIND <= SEG_SPHERE) {
            if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
                for (int st = 0; st < 4; ++st) {
                    if (f[st]) {
                        if (pd.count < PEND_K) {
                            pd.buf[pd.count * blockDim.x + threadIdx.x] = ((uint32_t)pos << 2) | (uint32_t)st;
                            ++pd.count;
                        } else if (pd.resume == NO_KEY) {
                            pd.resume = q * 4 + st;
                        }
                        t[st] = finf();
                    }
                }
            }
        }
#pragma unroll
        for (int st = 0; st < 4; ++st)
            if (t[st] < tb[st]) { tb[st] = t[st]; ib[st] = pos; }
    }
}

// FP64 stage of the MFMA form: entries (pos << 2) | set; the pair's ray is
// (set, lane % 16), regenerated from its index.
__device__ __forceinline__ void resolve_pending_mfma(const KParams &kp, CPTR(GeomDesc) G, const Emitter &E,
                                                     uint32_t gid, int emitter, int64_t wi0, int64_t base,
                                                     float (&tb)[4], int (&ib)[4], Pending &pd, uint32_t &nref)
{
    CPTR(rf_primitive) P = cst(G->prims);
    const int32_t *p2o = G->pos2orig;
    const int lane = threadIdx.x & 63;
    for (int e = 0; e < pd.count; ++e) {
        const uint32_t ent = pd.buf[e * blockDim.x + threadIdx.x];
        const int pos = (int)(ent >> 2);
        const int st = (int)(ent & 3u);
        const int orig = p2o[pos];
        if (orig < 0) continue;          // padding slot (never flags; kept out regardless)
        const int64_t rl = base + st * 16 + (lane & 15);
        float f[4], r4[4];
        workitem_floats(kp, gid, emitter, wi0 + (rl >> 2), f);
        rotate_rands(f, (int)(rl & 3), r4);
        const Ray64 r64 = widen(emit_fast(&E, r4));
        const rf_primitive p = ldc(P + orig);
        const float t = (float)hit_d_quadric(p.type, p.invm, p.util1, r64);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (st == i && (t < tb[i] || (t == tb[i] && pos < ib[i]))) { tb[i] = t; ib[i] = pos; }
    }
    nref += (uint32_t)pd.count;
    pd.count = 0;
}

// The wave's 64 rays (base + lane) as the MFMA B operands: lane l supplies
// component l / 16 of ray (set, l % 16) for origins ([s; 1]) and directions
// ([c; 0]).
__device__ __forceinline__ void mfma_rays(const KParams &kp, uint32_t gid, int emitter, int64_t wi0, int64_t rl,
                                          const Emitter &E, int lane, float (&Bo)[4], float (&Bd)[4])
{
    float f[4], r4[4];
    workitem_floats(kp, gid, emitter, wi0 + (rl >> 2), f);
    rotate_rands(f, (int)(rl & 3), r4);
    const Ray32 my = emit_fast(&E, r4);
    const int j = lane & 15, g = lane >> 4;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int src = st * 16 + j;
        const float vx = __shfl(my.sx, src), vy = __shfl(my.sy, src), vz = __shfl(my.sz, src);
        const float wx = __shfl(my.cx, src), wy = __shfl(my.cy, src), wz = __shfl(my.cz, src);
        Bo[st] = g == 0 ? vx : (g == 1 ? vy : (g == 2 ? vz : 1.0f));
        Bd[st] = g == 0 ? wx : (g == 1 ? wy : (g == 2 ? wz : 0.0f));
    }
}

template <bool REFINE, bool TRACE>
__device__ __forceinline__ void run_unit_mfma(const KParams &kp, CPTR(GeomDesc) G, uint32_t gid, int emitter,
                                              int64_t wi0, int64_t n_wi, uint32_t *bins, uint32_t *pend,
                                              uint32_t &nref, uint32_t &nover)
{
    const Emitter E = ldc(cst(G->emitters) + emitter);
    const int64_t ray0 = (wi0 - kp.trace_wi0) * 4;
    const float *Q = G->quads;
    const int32_t *p2o = G->pos2orig;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int64_t n_rays = n_wi * 4;
    for (int64_t base = (int64_t)wave * 64; base < n_rays; base += (int64_t)nwaves * 64) {
        const int64_t rl = base + lane;
        float Bo[4], Bd[4], tb[4];
        int ib[4];
        mfma_rays(kp, gid, emitter, wi0, rl, E, lane, Bo, Bd);
#pragma unroll
        for (int st = 0; st < 4; ++st) { tb[st] = finf(); ib[st] = -1; }
        Pending pd{pend, 0, NO_KEY, 0, vcopy(kp.tau), vcopy(EPS_F)};
        quad_loop<SEG_CYL, REFINE, false>(Q, G->qseg_begin[SEG_CYL], G->qseg_end[SEG_CYL], Bo, Bd, tb, ib, pd, lane);
        quad_loop<SEG_FRUSTUM, REFINE, false>(Q, G->qseg_begin[SEG_FRUSTUM], G->qseg_end[SEG_FRUSTUM], Bo, Bd, tb, ib,
                                              pd, lane);
        quad_loop<SEG_SPHERE, REFINE, false>(Q, G->qseg_begin[SEG_SPHERE], G->qseg_end[SEG_SPHERE], Bo, Bd, tb, ib, pd,
                                             lane);
        quad_loop<SEG_RING, REFINE, false>(Q, G->qseg_begin[SEG_RING], G->qseg_end[SEG_RING], Bo, Bd, tb, ib, pd, lane);
        quad_loop<SEG_RECT, REFINE, false>(Q, G->qseg_begin[SEG_RECT], G->qseg_end[SEG_RECT], Bo, Bd, tb, ib, pd, lane);
        if constexpr (REFINE) {
            for (;;) {
                if (__builtin_amdgcn_ballot_w64(pd.count > 0))
                    resolve_pending_mfma(kp, G, E, gid, emitter, wi0, base, tb, ib, pd, nref);
                if (!__builtin_amdgcn_ballot_w64(pd.resume != NO_KEY)) break;
                // a lane's list filled up: rebuild the operands, re-scan the
                // quadric quads from the first dropped pair
                nover += pd.resume != NO_KEY;
                pd.from = pd.resume;
                pd.resume = NO_KEY;
                float Bo2[4], Bd2[4];
                mfma_rays(kp, gid, emitter, wi0, rl + opaque_zero(), E, lane, Bo2, Bd2);
                const int lo = wave_min(pd.from) >> 2;
                const auto beg = [&](int s) { return G->qseg_begin[s] > lo ? G->qseg_begin[s] : lo; };
                quad_loop<SEG_CYL, true, true>(Q, beg(SEG_CYL), G->qseg_end[SEG_CYL], Bo2, Bd2, tb, ib, pd, lane);
                quad_loop<SEG_FRUSTUM, true, true>(Q, beg(SEG_FRUSTUM), G->qseg_end[SEG_FRUSTUM], Bo2, Bd2, tb, ib,
                                                   pd, lane);
                quad_loop<SEG_SPHERE, true, true>(Q, beg(SEG_SPHERE), G->qseg_end[SEG_SPHERE], Bo2, Bd2, tb, ib, pd,
                                                  lane);
            }
        }
        // nearest over the 4 primitive slots of each ray (lanes j, j+16, j+32, j+48)
        int mine = -1;
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            float t = tb[st];
            int p = ib[st];
#pragma unroll
            for (int off = 16; off <= 32; off <<= 1) {
                const float t2 = __shfl_xor(t, off);
                const int p2 = __shfl_xor(p, off);
                const bool take = (t2 < t) || (t2 == t && (unsigned)p2 < (unsigned)p);
                t = take ? t2 : t;
                p = take ? p2 : p;
            }
            if (g == st) mine = p;
        }
        if (rl < n_rays) {
            const int idx = mine >= 0 ? p2o[mine] : -1;
            tally<TRACE, false>(kp, G, emitter, ray0 + rl, idx, bins);
        }
    }
}

template <bool REFINE, bool TRACE>
__global__ void __launch_bounds__(256) k_queue_mfma(KParams kp)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
    uint32_t *bins = smem;
    uint32_t *pend = smem + (TRACE ? 0 : kp.bins);
    __shared__ int s_unit;
    uint32_t nref = 0, nover = 0;
    const bool lds_tally = !TRACE && kp.bins > 0;
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int)atomicAdd(kp.queue_head, 1u);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= kp.n_units) break;
        const UnitRef ur = unit_of(kp, u);
        CPTR(GeomDesc) G = ur.G;
        const int emitter = ur.emitter;
        const int64_t wi0 = ur.wi0, n_wi = ur.n_wi;
        const uint32_t gid = ur.gid;
        const int n = G->n;
        if (lds_tally) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) bins[i] = 0u;
            __syncthreads();
        }
        run_unit_mfma<REFINE, TRACE>(kp, G, gid, emitter, wi0, n_wi, bins, pend, nref, nover);
        if (lds_tally) {
            __syncthreads();
            flush_row(kp, ur, bins);
        }
        __syncthreads();
    }
    flush_stats(kp, 0, nref, nover);
}

// ---------------------------------------------------------------------------
// Fast BV form with rays binned by bounding surface: runBVGPU semantics
// (RayFactorGPU.cl:530-626) for scenes whose bounding surfaces do not ask for
// the intermediate check.  Per pass of blockDim x BNR x BSUB rays:
//   A. each thread emits BSUB x BNR rays and finds each ray's nearest
//      bounding surface (strict <, so the lowest BV index wins ties; flagged
//      BV roots are re-solved in FP64);
//   B. a counting sort in LDS buckets the rays by that surface;
//   C. every wave takes 64 x BNR rays of one bucket (buckets start on whole
//      waves) and scans that surface's children with the type-segmented fast
//      loop, BNR rays per lane.
// A ray that hits no bounding surface hits nothing, as in the reference.
// One ray per thread with divergent child ranges would scan the union of
// the wave's ranges; binned, a wave scans exactly one range.
constexpr int BNR = 4;
constexpr int BSUB = 4;          // sub-passes bucketed together: less padding per bucket
constexpr int WAVE_RAYS = 64 * BNR;

template <bool REFINE, bool TRACE>
__device__ __forceinline__ void run_unit_bvbin(const KParams &kp, CPTR(GeomDesc) G, uint32_t gid, int emitter,
                                               int64_t wi0, int64_t n_wi, uint32_t *bins, uint32_t *pend,
                                               uint32_t *ids, uint32_t *cnt, uint32_t *off, uint32_t *cur,
                                               uint32_t *pad, int *s_total, uint64_t &tests, uint64_t &cflops,
                                               uint32_t &nref, uint32_t &nover)
{
    const Emitter E = ldc(cst(G->emitters) + emitter);
    const int64_t ray0 = (wi0 - kp.trace_wi0) * 4;
    const int64_t n_rays = n_wi * 4;
    const int nbv = G->n_bvs;
    CPTR(FastPrim) FB = cst(G->fast_bvs);
    CPTR(rf_bounding_volume) B = cst(G->bvs);
    CPTR(BvMeta) M = cst(G->bv_meta);
    CPTR(FastPrim) FC = cst(G->bvc_fast);
    const float tau = vcopy(kp.tau), eps = vcopy(EPS_F);
    const int SP = (int)blockDim.x * BNR;            // rays per sub-pass
    const int P = SP * BSUB;                          // rays per pass (bucketed together)
    const int lid0 = (int)threadIdx.x * BNR;
    for (int64_t pass0 = 0; pass0 < n_rays; pass0 += P) {
        for (int b = threadIdx.x; b < nbv; b += blockDim.x) cnt[b] = 0u;
        __syncthreads();
        // A: nearest bounding surface of this thread's rays, BSUB sub-passes
        // (-2: beyond the unit)
        int bvi[BSUB * BNR];
#pragma unroll
        for (int sp = 0; sp < BSUB; ++sp) {
            const int lsp = sp * SP + lid0;
            Ray32 R[BNR];
            emit_group<BNR>(kp, gid, emitter, wi0, pass0 + lsp, E, R);
            float tbv[BNR];
            int bv[BNR];
#pragma unroll
            for (int i = 0; i < BNR; ++i) { tbv[i] = finf(); bv[i] = -1; }
            for (int w = 0; w < nbv; ++w) {
                const FastPrim q = vcopy(ldc(FB + w));
#pragma unroll
                for (int i = 0; i < BNR; ++i) {
                    bool f;
                    float t = fast_test_any<REFINE>(q, R[i], f, tau, eps);
                    if constexpr (REFINE) {
                        if (__builtin_amdgcn_ballot_w64(f)) {
                            if (f) {
                                // only quadrics flag
                                const rf_bounding_volume bvr = ldc(B + w);
                                t = (float)hit_d_quadric(bvr.type, bvr.invm, bvr.util1, widen(R[i]));
                                ++nref;
                            }
                        }
                    }
                    if (t < tbv[i]) { tbv[i] = t; bv[i] = w; }
                }
            }
#pragma unroll
            for (int i = 0; i < BNR; ++i) {
                if (pass0 + lsp + i >= n_rays) {
                    bv[i] = -2;
                } else {
                    tests += (uint64_t)nbv;
                    if (bv[i] >= 0) atomicAdd(&cnt[bv[i]], 1u);
                    else tally<TRACE, false>(kp, G, emitter, ray0 + pass0 + lsp + i, -1, bins);
                }
                bvi[sp * BNR + i] = bv[i];
            }
        }
        __syncthreads();
        // B: counting sort by bounding surface.  Rays are stored compactly by
        // bucket (off[b] = first slot of bucket b); each bucket also gets a
        // range of whole-wave slots (pad[b], a multiple of 64 x BNR), so that
        // every wave of phase C reads rays of one bucket only.
        if (threadIdx.x == 0) {
            uint32_t sum = 0, psum = 0;
            for (int b = 0; b < nbv; ++b) {
                const uint32_t c = cnt[b];
                off[b] = sum;
                cur[b] = sum;
                pad[b] = psum;
                sum += c;
                psum += (c + WAVE_RAYS - 1) / WAVE_RAYS * WAVE_RAYS;
            }
            *s_total = (int)psum;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BSUB * BNR; ++k) {
            if (bvi[k] >= 0) {
                const uint32_t at = atomicAdd(&cur[bvi[k]], 1u);
                ids[at] = (uint32_t)((k / BNR) * SP + lid0 + (k % BNR));
            }
        }
        __syncthreads();
        // C: each wave takes WAVE_RAYS consecutive padded slots (one bucket),
        // each lane BNR of them
        const int total = *s_total;
        const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
        for (int chunk0 = 0; chunk0 < total; chunk0 += SP) {
            const int s0 = chunk0 + wave * WAVE_RAYS;       // wave-uniform
            if (s0 >= total) continue;
            // bucket of this wave: last b with pad[b] <= s0 and a non-empty range
            int lo = 0, hi = nbv - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((int)pad[mid] <= s0) lo = mid; else hi = mid - 1;
            }
            while (lo > 0 && cnt[lo] == 0u) --lo;            // empty buckets share their pad offset
            const int b = __builtin_amdgcn_readfirstlane(lo);
            const int j0 = s0 - (int)pad[b] + lane * BNR;    // index within bucket b
            const int nb = (int)cnt[b];
            uint32_t lid[BNR];
            bool valid[BNR];
#pragma unroll
            for (int i = 0; i < BNR; ++i) {
                valid[i] = j0 + i < nb;
                lid[i] = valid[i] ? ids[off[b] + j0 + i] : 0u;
            }
            const uint32_t l0 = lid[0], l1 = lid[1], l2 = lid[2], l3 = lid[3];
            const auto ray_of = [=](int k) -> int64_t {
                const uint32_t l = k == 0 ? l0 : (k == 1 ? l1 : (k == 2 ? l2 : l3));
                return pass0 + (int64_t)l;
            };
            Ray32 R[BNR];
            float tb[BNR];
            int ib[BNR];
#pragma unroll
            for (int i = 0; i < BNR; ++i) {
                float f[4], r4[4];
                const int64_t rl = pass0 + lid[i];
                workitem_floats(kp, gid, emitter, wi0 + (rl >> 2), f);
                rotate_rands(f, (int)(rl & 3), r4);
                R[i] = emit_fast(&E, r4);
                tb[i] = finf();
                ib[i] = -1;
            }
            CPTR(int32_t) sg = cst(G->bvc_seg) + b * (N_SEG + 1);
            Pending pd{pend, 0, NO_KEY, 0, tau, eps};
            uint32_t csf = 0;     // the children's band stacks' flops (valid slots)
            uint32_t cev = 0;     // and their evaluations
            uint32_t vmask = 0;
#pragma unroll
            for (int i = 0; i < BNR; ++i) vmask |= valid[i] ? 1u << i : 0u;
            if (kp.stacks) {
                // the children's cylinder runs: band stacks share the quadric
                CPTR(int32_t) ro = cst(G->bvc_run_off) + b;
                const int j1 = ro[1];
                for (int j = ro[0]; j < j1; ++j) {
                    const CylRun run = ldc(cst(G->bvc_runs) + j);
                    if (kp.stacks >= 2 && run.stack >= 2) {
                        // searched in two slices of 2 rays (the search is
                        // compiled for 1 and 2); an overflow in either gives
                        // back every entry of the stack
                        static_assert(BNR == 4, "slices of 2 rays");
                        Ray32 (&R0)[2] = *reinterpret_cast<Ray32 (*)[2]>(&R[0]);
                        Ray32 (&R1)[2] = *reinterpret_cast<Ray32 (*)[2]>(&R[2]);
                        float (&tb0)[2] = *reinterpret_cast<float (*)[2]>(&tb[0]);
                        float (&tb1)[2] = *reinterpret_cast<float (*)[2]>(&tb[2]);
                        int (&ib0)[2] = *reinterpret_cast<int (*)[2]>(&ib[0]);
                        int (&ib1)[2] = *reinterpret_cast<int (*)[2]>(&ib[2]);
                        pd.stack0 = pd.count;
                        stack_loop<2, REFINE>(FC, cst(G->bvc_bands), run, run.stack, R0, tb0, ib0, pd, csf, cev,
                                              vmask & 3u);
                        pd.slot0 = 2;
                        stack_loop<2, REFINE>(FC, cst(G->bvc_bands), run, run.stack, R1, tb1, ib1, pd, csf, cev,
                                              vmask >> 2);
                        pd.slot0 = 0;
                        pd.stack0 = -1;
                    } else if (run.stack)
                        stack_loop<BNR, REFINE>(FC, cst(G->bvc_bands), run, 1, R, tb, ib, pd, csf, cev, vmask);
                    else
                        seg_loop<SEG_CYL, BNR, REFINE, false>(FC, run.begin, run.end, R, tb, ib, pd);
                }
            } else {
                seg_loop<SEG_CYL, BNR, REFINE, false>(FC, sg[SEG_CYL], sg[SEG_CYL + 1], R, tb, ib, pd);
            }
            seg_loop<SEG_FRUSTUM, BNR, REFINE, false>(FC, sg[SEG_FRUSTUM], sg[SEG_FRUSTUM + 1], R, tb, ib, pd);
            seg_loop<SEG_SPHERE, BNR, REFINE, false>(FC, sg[SEG_SPHERE], sg[SEG_SPHERE + 1], R, tb, ib, pd);
            seg_loop<SEG_RING, BNR, REFINE, false>(FC, sg[SEG_RING], sg[SEG_RING + 1], R, tb, ib, pd);
            seg_loop<SEG_RECT, BNR, REFINE, false>(FC, sg[SEG_RECT], sg[SEG_RECT + 1], R, tb, ib, pd);
            if constexpr (REFINE) {
                if (__builtin_amdgcn_ballot_w64(pd.count > 0))
                    resolve_pending<BNR>(kp, G, E, gid, emitter, wi0, ray_of, tb, ib, pd, nref);
                if (__builtin_amdgcn_ballot_w64(pd.resume != NO_KEY))
                    finish_dropped<BNR>(kp, G, E, gid, emitter, wi0, ray_of, FC, sg, sg + 1, tb, ib, pd, nref, nover);
            }
            const uint32_t cfl = cst(G->bv_cflops)[4 * b + (kp.stacks ? 1 : 0)];
            const uint32_t cnt1 = cst(G->bv_cflops)[4 * b + (kp.stacks ? 3 : 2)];
            cflops += csf;
            tests += cev;
#pragma unroll
            for (int i = 0; i < BNR; ++i) {
                if (valid[i]) {
                    tests += cnt1;
                    cflops += cfl;
                    tally<TRACE, false>(kp, G, emitter, ray0 + pass0 + lid[i], ib[i], bins);
                }
            }
        }
        __syncthreads();   // buckets are rebuilt by the next pass
    }
}

// Registers: 94 without the FP64 stage (5 waves per SIMD); with it the
// allocator wants ~170, capped at 128 (4 waves) with spills only outside the
// child-scan loops.
template <bool REFINE, bool TRACE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(REFINE ? 4 : 5))) k_queue_bvbin(KParams kp)
{
    // dynamic LDS: [bins][re-solve lists: PEND_K x blockDim][rays by bucket: blockDim x BNR x BSUB]
    //              [bucket counts, offsets, cursors, wave-padded offsets: 4 x bv_slots]
    extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
    uint32_t *bins = smem;
    uint32_t *pend = smem + (TRACE ? 0 : kp.bins);
    uint32_t *ids = pend + PEND_K * blockDim.x;
    uint32_t *cnt = ids + blockDim.x * BNR * BSUB;
    uint32_t *off = cnt + kp.bv_slots;
    uint32_t *cur = off + kp.bv_slots;
    uint32_t *pad = cur + kp.bv_slots;
    __shared__ int s_unit, s_total;
    uint64_t tests = 0, cflops = 0;
    uint32_t nref = 0, nover = 0;
    const bool lds_tally = !TRACE && kp.bins > 0;
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int)atomicAdd(kp.queue_head, 1u);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= kp.n_units) break;
        const UnitRef ur = unit_of(kp, u);
        CPTR(GeomDesc) G = ur.G;
        const int emitter = ur.emitter;
        const int64_t wi0 = ur.wi0, n_wi = ur.n_wi;
        const uint32_t gid = ur.gid;
        const int n = G->n;
        if (lds_tally) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) bins[i] = 0u;
            __syncthreads();
        }
        run_unit_bvbin<REFINE, TRACE>(kp, G, gid, emitter, wi0, n_wi, bins, pend, ids, cnt, off, cur, pad, &s_total,
                                      tests, cflops, nref, nover);
        if (lds_tally) {
            __syncthreads();
            flush_row(kp, ur, bins);
        }
        __syncthreads();
    }
    flush_stats(kp, tests, nref, nover, cflops);
}

// Persistent queue kernel: every workgroup pulls units until the queue is
// empty (all waves reach the same exit).
// Four rays per thread with the FP64 stage fits 96 VGPRs (five waves per
// SIMD) when asked to; left alone it takes a few more and drops to four.
// The compat scan (four rays per visit) is held to the same five waves.  Two
// rays per thread, uncounted, run at six (80 VGPRs, a few spilled outside the
// analytic band path): 3% faster on the 256-geometry sweep than five, 2.5%
// faster than seven; the counted variant stays at five (it spills heavily at
// six).
// COUNT: the fast band search counts its evaluations (off for timed launches:
// the counters cost ~2%, and a launch counted once gives the same counts).
// BT: each unit's band records (its geometry's cylinder segment) are copied
// to LDS, where the searched stacks read them.
// BVH: the units of scenes with a hierarchy run run_unit_bvh, the others the
// scan; 4 waves per SIMD (128 VGPRs: the walk keeps the next node's record
// and two leaf records in flight; the compiler's own choice, 136 VGPRs at 3
// waves, measured 20% slower).
template <int MATH, bool BV, bool TRACE, int NR, bool REFINE, bool COUNT = true, bool BT = false, bool BVH = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(BVH ? 4 : (NR == 2 && !COUNT) ? 6 : (NR == 4 || (MATH == KM_COMPAT && !BV)) ? 5 : (MATH == KM_FP64 ? 4 : 1)))) k_queue(KParams kp)
{
    // dynamic LDS: [bins: kp.bins x u32][pending re-solves: PEND_K x blockDim x u32]
    //              [BT: band records, 8 B each, from the next 8-byte boundary]
    //              [BVH: traversal stacks, (depth + 1) x blockDim x u32]
    extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
    uint32_t *bins = smem;
    uint32_t *pend = smem + (TRACE ? 0 : kp.bins);
    uint32_t *bstk = pend + PEND_K * blockDim.x + threadIdx.x;
    uint64_t *sband = (uint64_t *)(smem + ((kp.bins + PEND_K * blockDim.x + 1u) & ~1u));
    __shared__ int s_unit;
    uint64_t tests = 0, sflops = 0, sevals = 0;
    uint32_t nref = 0, nover = 0;
    const bool lds_tally = !TRACE && kp.bins > 0;
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int)atomicAdd(kp.queue_head, 1u);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= kp.n_units) break;
        const UnitRef ur = unit_of(kp, u);
        CPTR(GeomDesc) G = ur.G;
        const int emitter = ur.emitter;
        const int64_t wi0 = ur.wi0, n_wi = ur.n_wi;
        const uint32_t gid = ur.gid;
        const int n = G->n;
        BandsLds bl{};
        if constexpr (BT) {
            // the geometry's cylinder segment fits: the host sized the copy
            // for the largest segment of the launch
            const int cb = G->seg_begin[SEG_CYL], nc = G->seg_end[SEG_CYL] - cb;
            CPTR(uint64_t) src = (CPTR(uint64_t))(G->bands + cb);
            for (int i = threadIdx.x; i < nc; i += blockDim.x) sband[i] = src[i];
            bl = BandsLds{(LDS_PTR(uint64_t))sband, cb};
        }
        if (lds_tally)
            for (int i = threadIdx.x; i < n; i += blockDim.x) bins[i] = 0u;
        if (BT || lds_tally) __syncthreads();
        run_unit<MATH, BV, TRACE, NR, REFINE, false, COUNT, BT, BVH>(kp, G, gid, emitter, wi0, n_wi, threadIdx.x,
                                                                     blockDim.x, bins, pend, tests, sflops, sevals,
                                                                     nref, nover, bl, bstk);
        if (lds_tally) {
            __syncthreads();
            flush_row(kp, ur, bins);
        }
        __syncthreads();
    }
    flush_stats(kp, tests, nref, nover, sflops, sevals);
}

// Reference dispatch shape (RayFactorGPU.cl:481 launched as main.cpp:172-176):
// one work item per thread over the implicit unit, one global atomic per
// hitting ray.
template <int MATH, bool BV>
__global__ void __launch_bounds__(256) k_direct(KParams kp)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t smem[];   // pending re-solves: PEND_K x blockDim x u32
    CPTR(GeomDesc) G = cst(kp.geoms);
    uint64_t tests = 0;
    uint32_t nref = 0, nover = 0;
    const int64_t wl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t sflops = 0, sevals = 0;
    run_unit<MATH, BV, false, 1, true, true>(kp, G, kp.imp_geom_id, kp.imp_emitter, kp.imp_wi0, kp.imp_n_wi,
                                             wl, kp.imp_n_wi + 1, nullptr, smem, tests, sflops, sevals, nref, nover);
    flush_stats(kp, tests, nref, nover, sflops, sevals);
}

__global__ void k_fill_u32(unsigned int *p, unsigned int v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// FP64 normalisation of a hit matrix: vf = hits / rays (the int64 count is
// exact in double below 2^53; the division is IEEE double, as on the host).
__global__ void k_normalise(const unsigned long long *hits, double *vf, size_t n2, double rays)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) vf[i] = rays > 0.0 ? (double)hits[i] / rays : 0.0;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Start of a persistent solve in one launch: every geometry's n x n tally, the
// statistics and the queue counter to zero.  (One fill launch per geometry
// left gaps of about 0.2 ms per graph node between them: 30-55 ms on the
// 256-geometry sweep.)
__global__ void __launch_bounds__(256) k_fill_solve(const GeomDesc *geoms, int n_geoms, unsigned long long *stats,
                                                    unsigned int *queue_head)
{
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    if (tid < (size_t)kNStats) stats[tid] = 0ull;
    if (tid == 0 && queue_head) *queue_head = 0u;
    for (int g = 0; g < n_geoms; ++g) {
        CPTR(GeomDesc) d = cst(geoms + g);
        unsigned long long *h = d->hits;
        const size_t n2 = (size_t)d->n * (size_t)d->n;
        for (size_t i = tid; i < n2; i += nth) h[i] = 0ull;
    }
}

// ---------------------------------------------------------------------------
// Exhaustive check of the compat reciprocal / square root: every float bit
// pattern, each compared with the compiler's IEEE operation (bit for bit, or
// both NaN).  counts[k] = mismatches of check k, first[k] = lowest mismatching
// input bits:
//   0 FastOps::rcp where it reports ok, vs 1.0f / x            (must be 0)
//   1 FastOps::sqrt where it reports ok, vs sqrtf(x)            (must be 0)
//   2 recip_nr on every input vs 1.0f / x                       (information)
//   3 v_sqrt_f32 alone vs sqrtf(x), for normal x > 0            (information)
// ---------------------------------------------------------------------------
constexpr int SELFTEST_PER_THREAD = 256;

__device__ __forceinline__ bool same_bits(float a, float b)
{
    return __float_as_uint(a) == __float_as_uint(b) || (__builtin_isnan(a) && __builtin_isnan(b));
}

__global__ __launch_bounds__(256) void k_selftest_math(unsigned long long *counts, unsigned int *first)
{
    const uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * SELFTEST_PER_THREAD;
    uint32_t n[4] = {0, 0, 0, 0};
    uint32_t lo[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
    for (int i = 0; i < SELFTEST_PER_THREAD; ++i) {
        const uint32_t bits = (uint32_t)(base + i);
        const float x = __uint_as_float(bits);
        const float ri = recip_ieee(x), si = sqrt_ieee(x);
        bool bad[4];
        FastOps fr, fs;
        const float rf = fr.rcp(x), sf = fs.sqrt(x);
        bad[0] = fr.ok && !same_bits(rf, ri);
        bad[1] = fs.ok && !same_bits(sf, si);
        bad[2] = !same_bits(recip_nr(x), ri);
        bad[3] = __builtin_amdgcn_class(x, 0x100) && !same_bits(__builtin_amdgcn_sqrtf(x), si);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (bad[k]) { ++n[k]; lo[k] = bits < lo[k] ? bits : lo[k]; }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (n[k]) { atomicAdd(&counts[k], (unsigned long long)n[k]); atomicMin(&first[k], lo[k]); }
}

// Largest absolute error of the fast emission's sin and cos of 2 pi u over
// every u = k 2^-24, k < 2^24 (u01_24's values), against FP64 sin and cos:
// err[0] sin, err[1] cos, as the bits of non-negative floats (atomic max).
__global__ __launch_bounds__(256) void k_selftest_sincos(unsigned int *err)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const float u = (float)k * 0x1p-24f;
    const double ph = 6.283185307179586476925 * (double)u;
    float s, c;
    sincos_rev(u, &s, &c);
    const float es = (float)__builtin_fabs((double)s - sin(ph));
    const float ec = (float)__builtin_fabs((double)c - cos(ph));
    atomicMax(&err[0], __float_as_uint(es));
    atomicMax(&err[1], __float_as_uint(ec));
}

hipError_t launch_selftest_sincos(unsigned int *err, hipStream_t stream)
{
    // 2^24 inputs: 2^16 groups x 256 threads
    hipLaunchKernelGGL(k_selftest_sincos, dim3(1u << 16), dim3(256), 0, stream, err);
    return hipGetLastError();
}

hipError_t launch_selftest_math(unsigned long long *counts, unsigned int *first, hipStream_t stream)
{
    // 2^32 inputs: 2^16 groups x 256 threads x 256 inputs
    hipLaunchKernelGGL(k_selftest_math, dim3(1u << 16), dim3(256), 0, stream, counts, first);
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int *p, unsigned int value, size_t count, hipStream_t stream)
{
    if (!count) return hipSuccess;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, stream, p, value, count);
    return hipGetLastError();
}

hipError_t launch_normalise(const unsigned long long *hits, double *vf, size_t n2, double rays, hipStream_t stream)
{
    if (!n2) return hipSuccess;
    const unsigned blocks = (unsigned)((n2 + 255) / 256);
    hipLaunchKernelGGL(k_normalise, dim3(blocks), dim3(256), 0, stream, hits, vf, n2, rays);
    return hipGetLastError();
}

hipError_t launch_fill_solve(const GeomDesc *geoms, int n_geoms, size_t cells, unsigned long long *stats,
                             unsigned int *queue_head, hipStream_t stream)
{
    if (!geoms || !stats || n_geoms < 0) return hipErrorInvalidValue;
    const size_t blocks = std::min<size_t>(std::max<size_t>((cells + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_fill_solve, dim3((unsigned)blocks), dim3(256), 0, stream, geoms, n_geoms, stats, queue_head);
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
    if constexpr ((MATH == KM_FAST || MATH == KM_FP64) && !BV) {
        if (c.bvh) {
            bool count = true;
            if constexpr (!TRACE) count = c.count != 0;
            if (count)
                hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE, true, false, true>), dim3(c.grid),
                                   dim3(c.block), c.lds_bytes, s, kp);
            else if constexpr (!TRACE)
                hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE, false, false, true>), dim3(c.grid),
                                   dim3(c.block), c.lds_bytes, s, kp);
            return hipGetLastError();
        }
    }
    // the counter switch and the LDS band copy matter for the fast band search
    // only (1 and 2 rays per thread)
    if constexpr (MATH == KM_FAST && !BV && !TRACE && NR <= 2) {
        if (c.count) {
            if (c.bt) hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE, true, true>), dim3(c.grid), dim3(c.block),
                                         c.lds_bytes, s, kp);
            else hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE, true, false>), dim3(c.grid), dim3(c.block),
                                    c.lds_bytes, s, kp);
        } else {
            if (c.bt) hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE, false, true>), dim3(c.grid), dim3(c.block),
                                         c.lds_bytes, s, kp);
            else hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE, false, false>), dim3(c.grid), dim3(c.block),
                                    c.lds_bytes, s, kp);
        }
    } else {
        hipLaunchKernelGGL((k_queue<MATH, BV, TRACE, NR, REFINE>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
    }
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
    if constexpr (MATH == KM_FAST) {
        if (c.mfma && !c.use_bv && !c.direct) {
            if (c.trace) {
                if (c.refine) hipLaunchKernelGGL((k_queue_mfma<true, true>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
                else hipLaunchKernelGGL((k_queue_mfma<false, true>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
            } else {
                if (c.refine) hipLaunchKernelGGL((k_queue_mfma<true, false>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
                else hipLaunchKernelGGL((k_queue_mfma<false, false>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
            }
            return hipGetLastError();
        }
    }
    if constexpr (MATH == KM_FAST) {
        if (c.bvbin && c.use_bv && !c.direct) {
            if (c.trace) {
                if (c.refine) hipLaunchKernelGGL((k_queue_bvbin<true, true>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
                else hipLaunchKernelGGL((k_queue_bvbin<false, true>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
            } else {
                if (c.refine) hipLaunchKernelGGL((k_queue_bvbin<true, false>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
                else hipLaunchKernelGGL((k_queue_bvbin<false, false>), dim3(c.grid), dim3(c.block), c.lds_bytes, s, kp);
            }
            return hipGetLastError();
        }
    }
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
