// rf_device.hpp -- device arithmetic of the view-factor kernels (gfx950).
//
// Three arithmetic forms share the reference's sampling parametrisation and
// random-number use (RayFactorGPU.cl:481-527):
//   compat  FP32, operation-for-operation RayFactorGPU.cl with its
//           vendor-defined builtins pinned: native_recip -> IEEE 1/x,
//           native_sqrt -> IEEE sqrt, native_rsqrt -> 1/sqrt, mad -> fma,
//           sincos -> sincos_pinned (below), dot -> left-to-right sum of
//           products.  The translation unit is compiled with
//           -ffp-contract=off, so every fused multiply-add is an explicit
//           fmaf().
//   fp64    FP64 physics semantics: corrected sphere test, stable roots,
//           explicit negative-discriminant miss, Gram-Schmidt emission frame.
//   fast    FP32 physics semantics with hardware v_rcp_f32 / v_sqrt_f32,
//           per-primitive precomputed rows, and flags for roots that are
//           ill-conditioned in FP32 (re-solved in FP64 by the caller).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rayfactor.h"

namespace rfd {

constexpr float EPS_F = 1E-3f;                 // RayFactorGPU.cl:7
constexpr double EPS_D = 1E-3;
constexpr float M_2PI_F = 2.0f * 0x1.921fb6p+1f;  // RayFactorGPU.cl:8 (2 * M_PI_F)
constexpr double TWO_PI_D = 6.283185307179586476925;

__device__ __forceinline__ float finf() { return __builtin_inff(); }
__device__ __forceinline__ double dinf() { return __builtin_inf(); }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Random123/philox.h:206-305).  Each 32x32->64 product is one
// v_mad_u64_u32 (both halves at once) rather than a v_mul_hi_u32 and a
// v_mul_lo_u32: the integer multiplies are the costliest VALU operations of
// ray emission.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k.x += 0x9E3779B9u; k.y += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

// Random block of one 4-ray work item.  Reference: key {gid, 0xdecafbad},
// ctr {1, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d} (RayFactorGPU.cl:486-497).
// Seeded: key {wi_lo, seed}, ctr {wi_hi, emitter, geometry, 0xbeeff00d}.
__device__ __forceinline__ uint4 workitem_u32(int rng, uint32_t seed, uint32_t geometry,
                                              int emitter, int64_t wi)
{
    if (rng == RF_RNG_REFERENCE)
        return philox4x32_10(make_uint4(1u, 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du),
                             make_uint2((uint32_t)wi, 0xdecafbadu));
    return philox4x32_10(make_uint4((uint32_t)((uint64_t)wi >> 32), (uint32_t)emitter, geometry,
                                    0xbeeff00du),
                         make_uint2((uint32_t)wi, seed));
}

// u01_closed_open_4x32_24 (RayFactorGPU.cl:59-62): exact.
__device__ __forceinline__ float u01_24(uint32_t u) { return (float)(u >> 8) * 0x1p-24f; }

// ---------------------------------------------------------------------------
// Pinned sincos: Cody-Waite reduction by pi/2 (three-part constant, exact
// first step under FMA for |x| < 8), Taylor polynomials on [-pi/4, pi/4].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos_pinned(float x, float *s, float *c)
{
    const float kf = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-kf, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-kf, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-kf, -0x1.ee59dap-50f, r);
    const int k = (int)kf;
    const float z = r * r;
    float ps = __builtin_fmaf(z, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    ps = __builtin_fmaf(z, ps, 0x1.111112p-7f);
    ps = __builtin_fmaf(z, ps, -0x1.555556p-3f);
    const float sn = __builtin_fmaf(r * z, ps, r);
    float pc = __builtin_fmaf(z, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    pc = __builtin_fmaf(z, pc, -0x1.6c16c2p-10f);
    pc = __builtin_fmaf(z, pc, 0x1.555556p-5f);
    pc = __builtin_fmaf(z, pc, -0.5f);
    const float cs = __builtin_fmaf(z, pc, 1.0f);
    const int q = k & 3;
    const float a = (q & 1) ? cs : sn;
    const float b = (q & 1) ? sn : cs;
    *s = (q & 2) ? -a : a;
    *c = ((q + 1) & 2) ? -b : b;
}

__device__ __forceinline__ void sincos_pinned64(double x, double *s, double *c)
{
    const double kf = __builtin_rint(x * 0x1.45f306dc9c883p-1);
    double r = __builtin_fma(-kf, 0x1.921fb54442d18p+0, x);
    r = __builtin_fma(-kf, 0x1.1a62633145c07p-54, r);
    r = __builtin_fma(-kf, -0x1.f1976b7ed8fbcp-110, r);
    const int k = (int)kf;
    const double z = r * r;
    double ps = __builtin_fma(z, -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33);
    ps = __builtin_fma(z, ps, -0x1.ae64567f544e4p-26);
    ps = __builtin_fma(z, ps, 0x1.71de3a556c734p-19);
    ps = __builtin_fma(z, ps, -0x1.a01a01a01a01ap-13);
    ps = __builtin_fma(z, ps, 0x1.1111111111111p-7);
    ps = __builtin_fma(z, ps, -0x1.5555555555555p-3);
    const double sn = __builtin_fma(r * z, ps, r);
    double pc = __builtin_fma(z, 0x1.ae7f3e733b81fp-45, -0x1.93974a8c07c9dp-37);
    pc = __builtin_fma(z, pc, 0x1.1eed8eff8d898p-29);
    pc = __builtin_fma(z, pc, -0x1.27e4fb7789f5cp-22);
    pc = __builtin_fma(z, pc, 0x1.a01a01a01a01ap-16);
    pc = __builtin_fma(z, pc, -0x1.6c16c16c16c17p-10);
    pc = __builtin_fma(z, pc, 0x1.5555555555555p-5);
    pc = __builtin_fma(z, pc, -0.5);
    const double cs = __builtin_fma(z, pc, 1.0);
    const int q = k & 3;
    const double a = (q & 1) ? cs : sn;
    const double b = (q & 1) ? sn : cs;
    *s = (q & 2) ? -a : a;
    *c = ((q + 1) & 2) ? -b : b;
}

struct Ray32 { float sx, sy, sz, cx, cy, cz; };
struct Ray64 { double sx, sy, sz, cx, cy, cz; };

// ===========================================================================
// COMPAT (FP32, reference operation order)
// ===========================================================================
// native_recip / native_sqrt / native_rsqrt are pinned to IEEE 1/x and sqrt
// (correctly rounded).  The compiler's IEEE sequences cost 11 (division) and
// ~16 (square root) instructions.  Inside the primitive scan (find_c_multi)
// the hit tests use FastOps instead, which reach the same bits from the
// hardware approximations in 3 and 8 instructions:
//   recip_nr: v_rcp_f32 and one FMA Newton step, valid where the result is a
//     normal float;
//   sqrt_adj: v_sqrt_f32 and a one-ulp correction from the FMA residuals
//     x - s*(s -/+ ulp), valid for |x| >= 2^-96 (negative x and NaN give NaN).
// An operand outside those ranges clears FastOps::ok, and the scan repeats
// that ray's test with IeeeOps (a wave-uniform branch, rarely taken).
// rf_selftest_math checks "ok implies the same bits as 1.0f / x, sqrtf(x)"
// over all 2^32 inputs on the device.
__device__ __forceinline__ float recip_ieee(float x) { return 1.0f / x; }
__device__ __forceinline__ float sqrt_ieee(float x) { return __builtin_sqrtf(x); }

__device__ __forceinline__ float recip_nr(float x)
{
    const float r0 = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r0, 1.0f), r0, r0);
}

__device__ __forceinline__ float sqrt_adj(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = (__builtin_fmaf(-sd, s, x) <= 0.0f) ? sd : s;
    r = (__builtin_fmaf(-su, s, x) > 0.0f) ? su : r;
    return r;
}

// class mask of v_cmp_class_f32: -normal | +normal
constexpr int CLS_NORMAL = 0x108;
// sqrt_adj's residuals lose accuracy for tiny |x| (on gfx950 it differs from
// sqrtf for |x| < 2^-104, subnormals included); below this bound the IEEE
// operation runs
constexpr float SQRT_ADJ_MIN = 0x1p-96f;

struct IeeeOps {
    __device__ __forceinline__ float rcp(float x) { return recip_ieee(x); }
    __device__ __forceinline__ float sqrt(float x) { return sqrt_ieee(x); }
};

struct FastOps {
    bool ok = true;
    __device__ __forceinline__ float rcp(float x)
    {
        const float r = recip_nr(x);
        ok = ok & __builtin_amdgcn_class(r, CLS_NORMAL);
        return r;
    }
    __device__ __forceinline__ float sqrt(float x)
    {
        ok = ok & !(__builtin_fabsf(x) < SQRT_ADJ_MIN);
        return sqrt_adj(x);
    }
};

__device__ __forceinline__ float recip_c(float x) { return recip_ieee(x); }
__device__ __forceinline__ float sqrt_c(float x) { return sqrt_ieee(x); }
__device__ __forceinline__ float rsqrt_c(float x) { return 1.0f / sqrt_ieee(x); }

// xfrmRay (RayFactorGPU.cl:99-109)
__device__ __forceinline__ Ray32 xfrm_c(const float *__restrict__ m, const Ray32 &w)
{
    Ray32 o;
    o.sx = __builtin_fmaf(m[0], w.sx, __builtin_fmaf(m[1], w.sy, __builtin_fmaf(m[2], w.sz, m[3])));
    o.cx = __builtin_fmaf(m[0], w.cx, __builtin_fmaf(m[1], w.cy, m[2] * w.cz));
    o.sy = __builtin_fmaf(m[4], w.sx, __builtin_fmaf(m[5], w.sy, __builtin_fmaf(m[6], w.sz, m[7])));
    o.cy = __builtin_fmaf(m[4], w.cx, __builtin_fmaf(m[5], w.cy, m[6] * w.cz));
    o.sz = __builtin_fmaf(m[8], w.sx, __builtin_fmaf(m[9], w.sy, __builtin_fmaf(m[10], w.sz, m[11])));
    o.cz = __builtin_fmaf(m[8], w.cx, __builtin_fmaf(m[9], w.cy, m[10] * w.cz));
    return o;
}

// shared tail of cylinderHit (:199-221) and frustumHit (:237-259)
template <class Ops>
__device__ __forceinline__ float roots_select_c(float A, float B, float C, const Ray32 &o, Ops &ops)
{
    const float discrim = ops.sqrt(B * B - A * C);
    float t1, t2;
    t1 = t2 = ops.rcp(A);
    t1 *= (B - discrim);
    t2 *= (B + discrim);
    const float z1 = o.sz + o.cz * t1;
    const float z2 = o.sz + o.cz * t2;
    if (t1 < EPS_F || z1 < 0.f || z1 > 1.f) t1 = finf();
    if (t2 < t1 && z2 > 0.f && z2 < 1.0f && t2 > EPS_F) return t2;
    return t1;
}

// One test of findObject's switch (RayFactorGPU.cl:452-461), hit tests :126-291.
// One type of the compat hit test (operation for operation as the cited
// reference lines); hit_c dispatches on the type.
template <int TYPE, class Ops>
__device__ __forceinline__ float hit_c_ops(const float *__restrict__ m, float u1, float u2, const Ray32 &w, Ops &ops)
{
    const Ray32 o = xfrm_c(m, w);
    if constexpr (TYPE == RF_RECTANGLE) {                              // :126-145
        const float th = -o.sz * ops.rcp(o.cz);
        const float hx = __builtin_fabsf(o.sx + o.cx * th);
        const float hy = __builtin_fabsf(o.sy + o.cy * th);
        if (hx > 1.f || hy > 1.f || th < EPS_F) return finf();
        return th;
    } else if constexpr (TYPE == RF_DISC) {                            // :147-166
        const float th = -o.sz * ops.rcp(o.cz);
        const float hx = o.sx + o.cx * th;
        const float hy = o.sy + o.cy * th;
        const float e = hx * hx + hy * hy;
        if (e <= 1.f && th > EPS_F) return th;
        return finf();
    } else if constexpr (TYPE == RF_ANNULUS) {                         // :168-187
        const float th = -o.sz * ops.rcp(o.cz);
        const float hx = o.sx + o.cx * th;
        const float hy = o.sy + o.cy * th;
        const float e = __builtin_fmaf(hx, hx, hy * hy);
        if (e >= u1 && e <= u2 && th > EPS_F) return th;
        return finf();
    } else if constexpr (TYPE == RF_CYLINDER) {                        // :189-222
        const float A = __builtin_fmaf(o.cx, o.cx, o.cy * o.cy);
        const float B = -__builtin_fmaf(o.sx, o.cx, o.sy * o.cy);
        const float C = __builtin_fmaf(o.sx, o.sx, __builtin_fmaf(o.sy, o.sy, -1.f));
        return roots_select_c(A, B, C, o, ops);
    } else if constexpr (TYPE == RF_FRUSTUM) {                         // :224-260
        const float fDir = (u1 - 1.0f) * o.cz;
        const float fStart = (u1 - 1.0f) * o.sz + 1.0f;
        const float A = __builtin_fmaf(o.cx, o.cx, __builtin_fmaf(o.cy, o.cy, (-fDir) * fDir));
        const float B = -__builtin_fmaf(o.sx, o.cx, __builtin_fmaf(o.sy, o.cy, (-fStart) * fDir));
        const float C = __builtin_fmaf(o.sx, o.sx, __builtin_fmaf(o.sy, o.sy, (-fStart) * fStart));
        return roots_select_c(A, B, C, o, ops);
    } else {                                                           // :262-291 (no sqrt, as the reference)
        const float A = o.cx * o.cx + o.cy * o.cy + o.cz * o.cz;
        const float B = o.cx * o.sx + o.cy * o.sy + o.cz * o.sz;
        const float C = o.sx * o.sx + o.sy * o.sy + o.sz * o.sz - 1.0f;
        float discrim = B * B - A * C;
        discrim = (discrim < 0.0f) ? finf() : discrim;
        if (discrim > 0.f) {
            const float t1 = (-B - discrim) * ops.rcp(A);
            const float t2 = (-B + discrim) * ops.rcp(A);
            if (t2 < t1 && t2 > EPS_F) return t2;
            else if (t1 > EPS_F) return t1;
        }
        return finf();
    }
}

template <int TYPE>
__device__ __forceinline__ float hit_c_t(const float *__restrict__ m, float u1, float u2, const Ray32 &w)
{
    IeeeOps ops;
    return hit_c_ops<TYPE>(m, u1, u2, w, ops);
}

__device__ __forceinline__ float hit_c(int type, const float *__restrict__ m, float u1, float u2,
                                       const Ray32 &w)
{
    switch (type) {
    case RF_RECTANGLE: return hit_c_t<RF_RECTANGLE>(m, u1, u2, w);
    case RF_DISC: return hit_c_t<RF_DISC>(m, u1, u2, w);
    case RF_ANNULUS: return hit_c_t<RF_ANNULUS>(m, u1, u2, w);
    case RF_CYLINDER: return hit_c_t<RF_CYLINDER>(m, u1, u2, w);
    case RF_FRUSTUM: return hit_c_t<RF_FRUSTUM>(m, u1, u2, w);
    case RF_SPHERE: return hit_c_t<RF_SPHERE>(m, u1, u2, w);
    default: return finf();
    }
}

// transformNormal (:73-85)
__device__ __forceinline__ void normal_c(const float *__restrict__ iv, float nx, float ny, float nz,
                                         float *w)
{
    w[0] = __builtin_fmaf(iv[0], nx, __builtin_fmaf(iv[4], ny, iv[8] * nz));
    w[1] = __builtin_fmaf(iv[1], nx, __builtin_fmaf(iv[5], ny, iv[9] * nz));
    w[2] = __builtin_fmaf(iv[2], nx, __builtin_fmaf(iv[6], ny, iv[10] * nz));
    const float t = rsqrt_c(__builtin_fmaf(w[0], w[0], __builtin_fmaf(w[1], w[1], w[2] * w[2])));
    w[0] *= t; w[1] *= t; w[2] *= t;
}

// osRayDirection (:111-121) + alignRayDirection (:87-97)
__device__ __forceinline__ void direction_c(const float r4[4], const float n[3], Ray32 &ray)
{
    float o0, o1, o2;
    sincos_pinned(r4[2] * M_2PI_F, &o1, &o0);
    float t = r4[3];
    o2 = sqrt_c(1.0f - t);
    t = sqrt_c(t);
    o0 = t * o0;
    o1 = t * o1;
    float d = recip_c(1.0f + n[2]);
    if (__builtin_isinf(d)) d = 0.5f;
    ray.cx = (n[2] * o0 + (n[1] * d) * (n[1] * o0 - n[0] * o1)) + n[0] * o2;
    ray.cy = ((n[0] * d) * (n[0] * o1 - n[1] * o0) + n[2] * o1) + n[1] * o2;
    ray.cz = ((-n[0]) * o0 - n[1] * o1) + n[2] * o2;
}

// transformPoint (:64-71), dot pinned left to right.
__device__ __forceinline__ void point_c(const float *__restrict__ m, float x, float y, float z, Ray32 &ray)
{
    ray.sx = ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
    ray.sy = ((m[4] * x + m[5] * y) + m[6] * z) + m[7];
    ray.sz = ((m[8] * x + m[9] * y) + m[10] * z) + m[11];
}

// Emission samplers (:298-443)
__device__ __forceinline__ Ray32 emit_c(const rf_primitive *__restrict__ p, const float r4[4])
{
    Ray32 ray;
    float vx, vy, vz, n[3], sn, cs, t;
    const int type = p->type;
    const float u1 = p->util1, u2 = p->util2;
    if (type == RF_DISC || type == RF_ANNULUS) {                       // :298-345
        sincos_pinned(r4[0] * M_2PI_F, &sn, &cs);
        vx = sn; vy = cs;
        t = (type == RF_DISC) ? sqrt_c(r4[1]) : sqrt_c((u2 - u1) * r4[1] + u1);
        vx = t * vx; vy = t * vy;
        point_c(p->m, vx, vy, 0.f, ray);
        normal_c(p->invm, 0.f, 0.f, 1.f, n);
    } else if (type == RF_CYLINDER) {                                  // :347-365
        sincos_pinned(r4[0] * M_2PI_F, &sn, &cs);
        vy = sn; vx = cs; vz = r4[1];
        point_c(p->m, vx, vy, vz, ray);
        normal_c(p->invm, vx * u2, vy * u2, 0.f, n);
    } else if (type == RF_FRUSTUM) {                                   // :369-393
        vz = (sqrt_c((1.0f - r4[0]) + (r4[0] * u1) * u1) - 1.0f) * recip_c(u1 - 1.0f);
        t = 1.0f + vz * (u1 - 1.0f);
        sincos_pinned(M_2PI_F * r4[1], &sn, &cs);
        vy = sn; vx = cs;
        vx *= t; vy *= t;
        point_c(p->m, vx, vy, vz, ray);
        normal_c(p->invm, vx * u2, vy * u2, ((1.f - u1) * (1.f + (u1 - 1.f) * vz)) * u2, n);
    } else if (type == RF_RECTANGLE) {                                 // :399-417
        vx = 2.0f * r4[0] - 1.0f;
        vy = 2.0f * r4[1] - 1.0f;
        point_c(p->m, vx, vy, 0.f, ray);
        normal_c(p->invm, 0.f, 0.f, 1.f, n);
    } else {                                                           // sphere :419-443
        vz = 2.0f * r4[0] - 1.0f;
        t = sqrt_c(1.0f - vz * vz);
        sincos_pinned(M_2PI_F * r4[1], &sn, &cs);
        vy = sn; vx = cs;
        vx *= t; vy *= t;
        point_c(p->m, vx, vy, vz, ray);
        normal_c(p->invm, vx * u2, vy * u2, vz * u2, n);
    }
    direction_c(r4, n, ray);
    return ray;
}

// ===========================================================================
// FP64 physics semantics
// ===========================================================================
__device__ __forceinline__ Ray64 xfrm_d(const float *__restrict__ iv, const Ray64 &w)
{
    Ray64 o;
    const double a0 = iv[0], a1 = iv[1], a2 = iv[2], a3 = iv[3];
    const double b0 = iv[4], b1 = iv[5], b2 = iv[6], b3 = iv[7];
    const double c0 = iv[8], c1 = iv[9], c2 = iv[10], c3 = iv[11];
    o.sx = ((a0 * w.sx + a1 * w.sy) + a2 * w.sz) + a3;
    o.cx = (a0 * w.cx + a1 * w.cy) + a2 * w.cz;
    o.sy = ((b0 * w.sx + b1 * w.sy) + b2 * w.sz) + b3;
    o.cy = (b0 * w.cx + b1 * w.cy) + b2 * w.cz;
    o.sz = ((c0 * w.sx + c1 * w.sy) + c2 * w.sz) + c3;
    o.cz = (c0 * w.cx + c1 * w.cy) + c2 * w.cz;
    return o;
}

// Smallest root t > EPS of A t^2 + 2 B t + C = 0 with z in [0,1] (zcheck).
__device__ __forceinline__ double quad_select_d(double A, double B, double C, const Ray64 &o, bool zcheck)
{
    double r1, r2;
    if (A == 0.0) {
        if (B == 0.0) return dinf();
        r1 = r2 = -C / (2.0 * B);
    } else {
        const double D = B * B - A * C;
        if (D < 0.0) return dinf();
        const double sq = __builtin_sqrt(D);
        const double q = -(B + (B >= 0.0 ? sq : -sq));
        if (q == 0.0) return dinf();
        r1 = q / A;
        r2 = C / q;
    }
    const double lo = r1 < r2 ? r1 : r2;
    const double hi = r1 < r2 ? r2 : r1;
    if (lo > EPS_D) {
        const double z = o.sz + o.cz * lo;
        if (!zcheck || (z >= 0.0 && z <= 1.0)) return lo;
    }
    if (hi > EPS_D) {
        const double z = o.sz + o.cz * hi;
        if (!zcheck || (z >= 0.0 && z <= 1.0)) return hi;
    }
    return dinf();
}

__device__ __forceinline__ double hit_d(int type, const float *__restrict__ iv, float u1, float u2,
                                        const Ray64 &w)
{
    const Ray64 o = xfrm_d(iv, w);
    switch (type) {
    case RF_RECTANGLE: case RF_DISC: case RF_ANNULUS: {
        if (o.cz == 0.0) return dinf();
        const double t = -o.sz / o.cz;
        if (!(t > EPS_D)) return dinf();
        const double hx = o.sx + o.cx * t;
        const double hy = o.sy + o.cy * t;
        if (type == RF_RECTANGLE) return (__builtin_fabs(hx) <= 1.0 && __builtin_fabs(hy) <= 1.0) ? t : dinf();
        const double e = hx * hx + hy * hy;
        if (type == RF_DISC) return e <= 1.0 ? t : dinf();
        return (e >= (double)u1 && e <= (double)u2) ? t : dinf();
    }
    case RF_CYLINDER: {
        const double A = o.cx * o.cx + o.cy * o.cy;
        const double B = o.sx * o.cx + o.sy * o.cy;
        const double C = (o.sx * o.sx + o.sy * o.sy) - 1.0;
        return quad_select_d(A, B, C, o, true);
    }
    case RF_FRUSTUM: {
        const double k = (double)u1 - 1.0;
        const double ws = 1.0 + k * o.sz;
        const double wc = k * o.cz;
        const double A = (o.cx * o.cx + o.cy * o.cy) - wc * wc;
        const double B = (o.sx * o.cx + o.sy * o.cy) - ws * wc;
        const double C = (o.sx * o.sx + o.sy * o.sy) - ws * ws;
        return quad_select_d(A, B, C, o, true);
    }
    case RF_SPHERE: {
        const double A = (o.cx * o.cx + o.cy * o.cy) + o.cz * o.cz;
        const double B = (o.sx * o.cx + o.sy * o.cy) + o.sz * o.cz;
        const double C = ((o.sx * o.sx + o.sy * o.sy) + o.sz * o.sz) - 1.0;
        return quad_select_d(A, B, C, o, false);
    }
    default:
        return dinf();
    }
}

// FP64 re-solve of a flagged quadric pair (cylinder, frustum, sphere) in one
// straight-line form: the same A, B, C and root selection as hit_d for these
// types, without the plane branches (fewer live registers where it is inlined).
__device__ __forceinline__ double hit_d_quadric(int type, const float *__restrict__ iv, float u1, const Ray64 &w)
{
    const Ray64 o = xfrm_d(iv, w);
    const bool fr = type == RF_FRUSTUM, sp = type == RF_SPHERE;
    const double k = fr ? (double)u1 - 1.0 : 0.0;
    const double ws = 1.0 + k * o.sz;
    const double wc = k * o.cz;
    double A = o.cx * o.cx + o.cy * o.cy;
    double B = o.sx * o.cx + o.sy * o.cy;
    double C = o.sx * o.sx + o.sy * o.sy;
    if (sp) { A = A + o.cz * o.cz; B = B + o.sz * o.cz; C = (C + o.sz * o.sz) - 1.0; }
    else if (fr) { A = A - wc * wc; B = B - ws * wc; C = C - ws * ws; }
    else C = C - 1.0;
    return quad_select_d(A, B, C, o, !sp);
}

// Physics emission (FP64).  Point parametrisation as RayFactorGPU.cl:298-443
// (frustum z in the cancellation-free form); frame: n = normalize(invm^T
// n_os), tangent = the world image of a fixed object axis made orthogonal to
// n, c = o.x t1 + o.y (n x t1) + o.z n.
__device__ __forceinline__ Ray64 emit_d(const rf_primitive *__restrict__ p, const double r4[4])
{
    Ray64 ray;
    double sn, cs, t, v0, v1, v2, n0, n1, n2;
    sincos_pinned64(r4[2] * TWO_PI_D, &sn, &cs);
    t = __builtin_sqrt(r4[3]);
    const double o0 = t * cs, o1 = t * sn, o2 = __builtin_sqrt(1.0 - r4[3]);
    const int type = p->type;
    const double u1 = p->util1, u2 = p->util2;
    if (type == RF_DISC || type == RF_ANNULUS) {
        sincos_pinned64(r4[0] * TWO_PI_D, &sn, &cs);
        t = (type == RF_DISC) ? __builtin_sqrt(r4[1]) : __builtin_sqrt((u2 - u1) * r4[1] + u1);
        v0 = t * sn; v1 = t * cs; v2 = 0.0;
        n0 = 0.0; n1 = 0.0; n2 = 1.0;
    } else if (type == RF_CYLINDER) {
        sincos_pinned64(r4[0] * TWO_PI_D, &sn, &cs);
        v0 = cs; v1 = sn; v2 = r4[1];
        n0 = cs * u2; n1 = sn * u2; n2 = 0.0;
    } else if (type == RF_FRUSTUM) {
        const double a = r4[0] * (u1 * u1 - 1.0);
        v2 = (r4[0] * (u1 + 1.0)) / (__builtin_sqrt(1.0 + a) + 1.0);
        const double rho = 1.0 + v2 * (u1 - 1.0);
        sincos_pinned64(r4[1] * TWO_PI_D, &sn, &cs);
        v0 = rho * cs; v1 = rho * sn;
        n0 = v0 * u2; n1 = v1 * u2; n2 = ((1.0 - u1) * rho) * u2;
    } else if (type == RF_RECTANGLE) {
        v0 = 2.0 * r4[0] - 1.0; v1 = 2.0 * r4[1] - 1.0; v2 = 0.0;
        n0 = 0.0; n1 = 0.0; n2 = 1.0;
    } else {
        v2 = 2.0 * r4[0] - 1.0;
        const double rho = __builtin_sqrt((1.0 - v2) * (1.0 + v2));
        sincos_pinned64(r4[1] * TWO_PI_D, &sn, &cs);
        v0 = rho * cs; v1 = rho * sn;
        n0 = v0 * u2; n1 = v1 * u2; n2 = v2 * u2;
    }
    const float *m = p->m, *iv = p->invm;
    ray.sx = (((double)m[0] * v0 + (double)m[1] * v1) + (double)m[2] * v2) + (double)m[3];
    ray.sy = (((double)m[4] * v0 + (double)m[5] * v1) + (double)m[6] * v2) + (double)m[7];
    ray.sz = (((double)m[8] * v0 + (double)m[9] * v1) + (double)m[10] * v2) + (double)m[11];
    double w0 = ((double)iv[0] * n0 + (double)iv[4] * n1) + (double)iv[8] * n2;
    double w1 = ((double)iv[1] * n0 + (double)iv[5] * n1) + (double)iv[9] * n2;
    double w2 = ((double)iv[2] * n0 + (double)iv[6] * n1) + (double)iv[10] * n2;
    double inv = 1.0 / __builtin_sqrt((w0 * w0 + w1 * w1) + w2 * w2);
    w0 *= inv; w1 *= inv; w2 *= inv;
    int k;
    if (type == RF_CYLINDER || type == RF_FRUSTUM) k = 2;
    else if (type == RF_SPHERE) {
        const double a0 = __builtin_fabs(n0), a1 = __builtin_fabs(n1), a2 = __builtin_fabs(n2);
        k = (a0 <= a1 && a0 <= a2) ? 0 : (a1 <= a2 ? 1 : 2);
    } else k = 0;
    // world image of object axis k (column k of m), chosen without indexing
    const double a0 = k == 0 ? m[0] : (k == 1 ? m[1] : m[2]);
    const double a1 = k == 0 ? m[4] : (k == 1 ? m[5] : m[6]);
    const double a2 = k == 0 ? m[8] : (k == 1 ? m[9] : m[10]);
    const double an = (a0 * w0 + a1 * w1) + a2 * w2;
    double t0 = a0 - an * w0, t1 = a1 - an * w1, t2 = a2 - an * w2;
    inv = 1.0 / __builtin_sqrt((t0 * t0 + t1 * t1) + t2 * t2);
    t0 *= inv; t1 *= inv; t2 *= inv;
    const double b0 = w1 * t2 - w2 * t1;
    const double b1 = w2 * t0 - w0 * t2;
    const double b2 = w0 * t1 - w1 * t0;
    ray.cx = (o0 * t0 + o1 * b0) + o2 * w0;
    ray.cy = (o0 * t1 + o1 * b1) + o2 * w1;
    ray.cz = (o0 * t2 + o1 * b2) + o2 * w2;
    return ray;
}

// ===========================================================================
// FAST (FP32 physics semantics, hardware reciprocal / square root)
// ===========================================================================
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq_fast(float x) { return __builtin_amdgcn_rsqf(x); }

// sin(2 pi u) and cos(2 pi u) for the fast emission: v_sin_f32 / v_cos_f32
// take their argument in revolutions, so u (u01_24's values in [0, 1)) needs
// no reduction.  Largest absolute error over all 2^24 such u, against FP64:
// 1.25e-7 for both (rf_selftest_sincos); the FP32 polynomial it replaced
// measured 8.3e-8.
__device__ __forceinline__ void sincos_rev(float u, float *s, float *c)
{
    *s = __builtin_amdgcn_sinf(u);
    *c = __builtin_amdgcn_cosf(u);
}

// Emission frame of a cylinder emitter, built once on the host in FP64
// (kind 1).  With A, B the first two rows of invm's linear part and a the
// axis (m's third column), the world normal at point angle phi is
// u2 (cos A + sin B) normalised and the tangent is a orthogonalised against
// it.  When A and B are orthogonal, of equal length and both orthogonal to a
// (no shear, a uniform cross-section scale), the normal is
// cos na + sin nb (na, nb = sign(u2) A/|A|, sign(u2) B/|B|), the tangent the
// unit axis ax, and the binormal (normal x tangent) cos pa + sin pb with
// pa = na x ax, pb = nb x ax.  Kind 0: the frame is built per ray.
struct EmitFrame {
    int32_t kind;
    float ax[3], na[3], nb[3], pa[3], pb[3];
};

// An emitter: its primitive and its frame (64 bytes after the 112).
struct Emitter : rf_primitive {
    EmitFrame f;
};

// Emission with the FP64 form's formulas, in FP32.
__device__ __forceinline__ Ray32 emit_fast(const rf_primitive *__restrict__ p, const float r4[4])
{
    Ray32 ray;
    float sn, cs, t, v0, v1, v2, n0, n1, n2;
    sincos_rev(r4[2], &sn, &cs);
    t = sqrt_fast(r4[3]);
    const float o0 = t * cs, o1 = t * sn, o2 = sqrt_fast(1.0f - r4[3]);
    const int type = p->type;
    const float u1 = p->util1, u2 = p->util2;
    if (type == RF_DISC || type == RF_ANNULUS) {
        sincos_rev(r4[0], &sn, &cs);
        t = (type == RF_DISC) ? sqrt_fast(r4[1]) : sqrt_fast(__builtin_fmaf(u2 - u1, r4[1], u1));
        v0 = t * sn; v1 = t * cs; v2 = 0.f;
        n0 = 0.f; n1 = 0.f; n2 = 1.f;
    } else if (type == RF_CYLINDER) {
        sincos_rev(r4[0], &sn, &cs);
        v0 = cs; v1 = sn; v2 = r4[1];
        n0 = cs * u2; n1 = sn * u2; n2 = 0.f;
    } else if (type == RF_FRUSTUM) {
        const float a = r4[0] * __builtin_fmaf(u1, u1, -1.0f);
        v2 = (r4[0] * (u1 + 1.0f)) * rcp_fast(sqrt_fast(1.0f + a) + 1.0f);
        const float rho = __builtin_fmaf(v2, u1 - 1.0f, 1.0f);
        sincos_rev(r4[1], &sn, &cs);
        v0 = rho * cs; v1 = rho * sn;
        n0 = v0 * u2; n1 = v1 * u2; n2 = ((1.0f - u1) * rho) * u2;
    } else if (type == RF_RECTANGLE) {
        v0 = __builtin_fmaf(2.0f, r4[0], -1.0f); v1 = __builtin_fmaf(2.0f, r4[1], -1.0f); v2 = 0.f;
        n0 = 0.f; n1 = 0.f; n2 = 1.f;
    } else {
        v2 = __builtin_fmaf(2.0f, r4[0], -1.0f);
        const float rho = sqrt_fast((1.0f - v2) * (1.0f + v2));
        sincos_rev(r4[1], &sn, &cs);
        v0 = rho * cs; v1 = rho * sn;
        n0 = v0 * u2; n1 = v1 * u2; n2 = v2 * u2;
    }
    const float *m = p->m, *iv = p->invm;
    ray.sx = __builtin_fmaf(m[0], v0, __builtin_fmaf(m[1], v1, __builtin_fmaf(m[2], v2, m[3])));
    ray.sy = __builtin_fmaf(m[4], v0, __builtin_fmaf(m[5], v1, __builtin_fmaf(m[6], v2, m[7])));
    ray.sz = __builtin_fmaf(m[8], v0, __builtin_fmaf(m[9], v1, __builtin_fmaf(m[10], v2, m[11])));
    float w0 = __builtin_fmaf(iv[0], n0, __builtin_fmaf(iv[4], n1, iv[8] * n2));
    float w1 = __builtin_fmaf(iv[1], n0, __builtin_fmaf(iv[5], n1, iv[9] * n2));
    float w2 = __builtin_fmaf(iv[2], n0, __builtin_fmaf(iv[6], n1, iv[10] * n2));
    float inv = rsq_fast(__builtin_fmaf(w0, w0, __builtin_fmaf(w1, w1, w2 * w2)));
    w0 *= inv; w1 *= inv; w2 *= inv;
    int k;
    if (type == RF_CYLINDER || type == RF_FRUSTUM) k = 2;
    else if (type == RF_SPHERE) {
        const float a0 = __builtin_fabsf(n0), a1 = __builtin_fabsf(n1), a2 = __builtin_fabsf(n2);
        k = (a0 <= a1 && a0 <= a2) ? 0 : (a1 <= a2 ? 1 : 2);
    } else k = 0;
    const float a0 = k == 0 ? m[0] : (k == 1 ? m[1] : m[2]);
    const float a1 = k == 0 ? m[4] : (k == 1 ? m[5] : m[6]);
    const float a2 = k == 0 ? m[8] : (k == 1 ? m[9] : m[10]);
    const float an = __builtin_fmaf(a0, w0, __builtin_fmaf(a1, w1, a2 * w2));
    float t0 = __builtin_fmaf(-an, w0, a0), t1 = __builtin_fmaf(-an, w1, a1), t2 = __builtin_fmaf(-an, w2, a2);
    inv = rsq_fast(__builtin_fmaf(t0, t0, __builtin_fmaf(t1, t1, t2 * t2)));
    t0 *= inv; t1 *= inv; t2 *= inv;
    const float b0 = __builtin_fmaf(w1, t2, -(w2 * t1));
    const float b1 = __builtin_fmaf(w2, t0, -(w0 * t2));
    const float b2 = __builtin_fmaf(w0, t1, -(w1 * t0));
    ray.cx = __builtin_fmaf(o0, t0, __builtin_fmaf(o1, b0, o2 * w0));
    ray.cy = __builtin_fmaf(o0, t1, __builtin_fmaf(o1, b1, o2 * w1));
    ray.cz = __builtin_fmaf(o0, t2, __builtin_fmaf(o1, b2, o2 * w2));
    return ray;
}

// The same with the emitter's frame: a cylinder with a kind-1 frame takes its
// normal, tangent and binormal from the frame's constants (the sampling,
// the point and the local direction are emit_fast's).
__device__ __forceinline__ Ray32 emit_fast(const Emitter *__restrict__ e, const float r4[4])
{
    if (e->f.kind != 1) return emit_fast(static_cast<const rf_primitive *>(e), r4);
    float s0, c0, s2, c2;
    sincos_rev(r4[0], &s0, &c0);
    sincos_rev(r4[2], &s2, &c2);
    const float t = sqrt_fast(r4[3]);
    const float o0 = t * c2, o1 = t * s2, o2 = sqrt_fast(1.0f - r4[3]);
    const float *m = e->m;
    const EmitFrame &f = e->f;
    Ray32 ray;
    ray.sx = __builtin_fmaf(m[0], c0, __builtin_fmaf(m[1], s0, __builtin_fmaf(m[2], r4[1], m[3])));
    ray.sy = __builtin_fmaf(m[4], c0, __builtin_fmaf(m[5], s0, __builtin_fmaf(m[6], r4[1], m[7])));
    ray.sz = __builtin_fmaf(m[8], c0, __builtin_fmaf(m[9], s0, __builtin_fmaf(m[10], r4[1], m[11])));
    float c[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float u = __builtin_fmaf(o1, f.pa[j], o2 * f.na[j]);
        const float v = __builtin_fmaf(o1, f.pb[j], o2 * f.nb[j]);
        c[j] = __builtin_fmaf(o0, f.ax[j], __builtin_fmaf(c0, u, s0 * v));
    }
    ray.cx = c[0];
    ray.cy = c[1];
    ray.cz = c[2];
    return ray;
}

// Fast-path primitive record (64 bytes, one s_load_dwordx16):
//   r[12]  rows of the world->object transform, with per-type folding:
//            cylinder  row 2 offset by -0.5 (accept |z'| <= 1/2)
//            frustum   row 2 replaced by w = 1 + (u1-1) z (cone radius)
//   k0,k1  ring: r^2 range (disc [0,1]); frustum: centre / half-width of w over z in [0,1]
//   orig   index of the primitive in the caller's array
//   type   rf_primitive_type
struct FastPrim {
    float r[12];
    float k0, k1;
    int32_t orig;
    int32_t type;
};

// Default relative margin tau of the FP32 conditioning flags; a flagged pair
// is re-solved in FP64.  The one margin guards three quantities, each several
// times (2^10) its FP32 error (transform + quadric in single precision):
//   |D| <= tau * b^2 with the tangent point ahead of the ray and on the band
//       (near-tangent: the root separation is ill-conditioned);
//   a root ahead of the ray within tau of a rim (band-height units);
//   frustum: |A| <= tau * (cx^2 + cy^2) (ray nearly along a generator).
constexpr float TAU_DEFAULT = 0x1p-14f;

}  // namespace rfd
