/*
 * rf_oracle.c -- CPU restatement of RayFactorCL/RayFactorGPU.cl.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle that the HIP path is
 * checked against, and the CPU baseline of the benchmark.  The product
 * library never links it.  Build with -ffp-contract=off (no contraction);
 * every fused multiply-add below is an explicit fma()/fmaf() call.
 *
 * Conventions pinned for the reference's vendor-defined builtins are listed
 * in rf_oracle.h.  Each function cites the reference file:line it follows.
 */
#include "rf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define EPS_F 1E-3f                       /* RayFactorGPU.cl:7 */
#define EPS_D 1E-3                        /* same constant, FP64 mode */
#define M_PI_F_ 0x1.921fb6p+1f           /* OpenCL M_PI_F (nearest float to pi) */
static const float M_2PI_F = 2.0f * M_PI_F_;   /* RayFactorGPU.cl:8 */
static const double TWO_PI_D = 6.283185307179586476925;

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Random123/philox.h:206-237, :259-305)                */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void rfo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }      /* bumpkey :270-275 */
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;                /* mulhilo32 :67-72,161 */
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;                                    /* round :259-268 */
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* RayFactorGPU.cl:59-62: (u >> 8) * 2^-24, exact. */
float rfo_u01_24(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

/* Reference key/counter: RayFactorGPU.cl:486-497 (and :546-554).
 * Seeded: key = {wi_lo, seed}, ctr = {wi_hi, emitter, geometry, 0xbeeff00d}. */
void rfo_workitem_u32(int32_t rng, uint32_t seed, uint32_t geometry, int32_t emitter,
                      int64_t wi, uint32_t out[4])
{
    uint32_t key[2], ctr[4];
    if (rng == RFO_RNG_REFERENCE) {
        key[0] = (uint32_t)wi; key[1] = 0xdecafbadu;
        ctr[0] = 1u; ctr[1] = 0xf00dcafeu; ctr[2] = 0xdeadbeefu; ctr[3] = 0xbeeff00du;
    } else {
        key[0] = (uint32_t)wi; key[1] = seed;
        ctr[0] = (uint32_t)((uint64_t)wi >> 32); ctr[1] = (uint32_t)emitter;
        ctr[2] = geometry; ctr[3] = 0xbeeff00du;
    }
    rfo_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------ */
/* Pinned sincos (stands for OpenCL sincos(), RayFactorGPU.cl:113 etc.) */
/* Cody-Waite reduction by pi/2 with a three-part constant, then the   */
/* Taylor polynomials of sin and cos on [-pi/4, pi/4].                 */
/* ------------------------------------------------------------------ */
void rfo_sincosf(float x, float *s, float *c)
{
    const float TWO_OVER_PI = 0x1.45f306p-1f;
    const float P1 = 0x1.921fb6p+0f;     /* pi/2 = P1 + P2 + P3 (72 bits) */
    const float P2 = -0x1.777a5cp-25f;
    const float P3 = -0x1.ee59dap-50f;
    const float kf = rintf(x * TWO_OVER_PI);
    float r = fmaf(-kf, P1, x);          /* exact for |x| < 2^3 */
    r = fmaf(-kf, P2, r);
    r = fmaf(-kf, P3, r);
    const int k = (int)kf;
    const float z = r * r;
    /* sin r = r + r^3 (S3 + z (S5 + z (S7 + z S9))), Sn = (-1)^((n-1)/2) / n! */
    float ps = fmaf(z, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    ps = fmaf(z, ps, 0x1.111112p-7f);
    ps = fmaf(z, ps, -0x1.555556p-3f);
    const float sn = fmaf(r * z, ps, r);
    /* cos r = 1 + z (C2 + z (C4 + z (C6 + z (C8 + z C10)))), Cn = (-1)^(n/2) / n! */
    float pc = fmaf(z, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    pc = fmaf(z, pc, -0x1.6c16c2p-10f);
    pc = fmaf(z, pc, 0x1.555556p-5f);
    pc = fmaf(z, pc, -0.5f);
    const float cs = fmaf(z, pc, 1.0f);
    switch (k & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

void rfo_sincos64(double x, double *s, double *c)
{
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb54442d18p+0;   /* pi/2 = P1 + P2 + P3 */
    const double P2 = 0x1.1a62633145c07p-54;
    const double P3 = -0x1.f1976b7ed8fbcp-110;
    const double kf = rint(x * TWO_OVER_PI);
    double r = fma(-kf, P1, x);
    r = fma(-kf, P2, r);
    r = fma(-kf, P3, r);
    const long k = (long)kf;
    const double z = r * r;
    /* Taylor to r^15 (sin) and r^16 (cos) on [-pi/4, pi/4] */
    double ps = fma(z, -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33);
    ps = fma(z, ps, -0x1.ae64567f544e4p-26);
    ps = fma(z, ps, 0x1.71de3a556c734p-19);
    ps = fma(z, ps, -0x1.a01a01a01a01ap-13);
    ps = fma(z, ps, 0x1.1111111111111p-7);
    ps = fma(z, ps, -0x1.5555555555555p-3);
    const double sn = fma(r * z, ps, r);
    double pc = fma(z, 0x1.ae7f3e733b81fp-45, -0x1.93974a8c07c9dp-37);
    pc = fma(z, pc, 0x1.1eed8eff8d898p-29);
    pc = fma(z, pc, -0x1.27e4fb7789f5cp-22);
    pc = fma(z, pc, 0x1.a01a01a01a01ap-16);
    pc = fma(z, pc, -0x1.6c16c16c16c17p-10);
    pc = fma(z, pc, 0x1.5555555555555p-5);
    pc = fma(z, pc, -0.5);
    const double cs = fma(z, pc, 1.0);
    switch (k & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

/* ------------------------------------------------------------------ */
/* COMPAT (FP32) helpers: RayFactorGPU.cl:64-121                        */
/* ------------------------------------------------------------------ */
static float recipf_(float x) { return 1.0f / x; }          /* native_recip */
static float rsqrtf_(float x) { return 1.0f / sqrtf(x); }   /* native_rsqrt */

/* transformPoint :64-71 (dot pinned to left-to-right sum of products). */
static void transform_point_f(const float *m, const float p[3], float w[3])
{
    for (int r = 0; r < 3; ++r) {
        const float *row = m + 4 * r;
        w[r] = ((row[0] * p[0] + row[1] * p[1]) + row[2] * p[2]) + row[3];
    }
}

/* transformNormal :73-85 */
static void transform_normal_f(const float *invm, const float n[3], float w[3])
{
    w[0] = fmaf(invm[0], n[0], fmaf(invm[4], n[1], invm[8] * n[2]));
    w[1] = fmaf(invm[1], n[0], fmaf(invm[5], n[1], invm[9] * n[2]));
    w[2] = fmaf(invm[2], n[0], fmaf(invm[6], n[1], invm[10] * n[2]));
    const float t = rsqrtf_(fmaf(w[0], w[0], fmaf(w[1], w[1], w[2] * w[2])));
    w[0] *= t; w[1] *= t; w[2] *= t;
}

/* alignRayDirection :87-97 */
static void align_ray_direction_f(float dir[3], const float o[3], const float n[3])
{
    float d = recipf_(1.0f + n[2]);
    if (isinf(d)) d = 0.5f;
    dir[0] = (n[2] * o[0] + (n[1] * d) * (n[1] * o[0] - n[0] * o[1])) + n[0] * o[2];
    dir[1] = ((n[0] * d) * (n[0] * o[1] - n[1] * o[0]) + n[2] * o[1]) + n[1] * o[2];
    dir[2] = ((-n[0]) * o[0] - n[1] * o[1]) + n[2] * o[2];
}

/* xfrmRay :99-109; ray = {s, c}, out likewise in object space. */
static void xfrm_ray_f(const float *invm, const float ray[6], float os[6])
{
    for (int r = 0; r < 3; ++r) {
        const float *m = invm + 4 * r;
        os[r] = fmaf(m[0], ray[0], fmaf(m[1], ray[1], fmaf(m[2], ray[2], m[3])));
        os[3 + r] = fmaf(m[0], ray[3], fmaf(m[1], ray[4], m[2] * ray[5]));
    }
}

/* osRayDirection :111-121 */
static void os_ray_direction_f(const float r4[4], float o[3])
{
    float sn, cs;
    rfo_sincosf(r4[2] * M_2PI_F, &sn, &cs);
    o[1] = sn; o[0] = cs;
    float t = r4[3];
    o[2] = sqrtf(1.0f - t);
    t = sqrtf(t);
    o[0] = t * o[0];
    o[1] = t * o[1];
}

/* ------------------------------------------------------------------ */
/* COMPAT hit tests :126-291                                           */
/* ------------------------------------------------------------------ */
static float rectangle_hit_f(const float *invm, const float ray[6])     /* :126-145 */
{
    float os[6]; xfrm_ray_f(invm, ray, os);
    const float th = -os[2] * recipf_(os[5]);
    const float hx = fabsf(os[0] + os[3] * th);
    const float hy = fabsf(os[1] + os[4] * th);
    if (hx > 1.f || hy > 1.f || th < EPS_F) return INFINITY;
    return th;
}

static float disc_hit_f(const float *invm, const float ray[6])          /* :147-166 */
{
    float os[6]; xfrm_ray_f(invm, ray, os);
    const float th = -os[2] * recipf_(os[5]);
    const float hx = os[0] + os[3] * th;
    const float hy = os[1] + os[4] * th;
    const float e = hx * hx + hy * hy;
    if (e <= 1.f && th > EPS_F) return th;
    return INFINITY;
}

static float annulus_hit_f(const float *invm, const float ray[6], float id2, float od2) /* :168-187 */
{
    float os[6]; xfrm_ray_f(invm, ray, os);
    const float th = -os[2] * recipf_(os[5]);
    const float hx = os[0] + os[3] * th;
    const float hy = os[1] + os[4] * th;
    const float e = fmaf(hx, hx, hy * hy);
    if (e >= id2 && e <= od2 && th > EPS_F) return th;
    return INFINITY;
}

static float quad_roots_select_f(float A, float B, float C, const float os[6])
{
    /* shared tail of cylinderHit :199-221 and frustumHit :237-259 */
    const float discrim = sqrtf(B * B - A * C);
    float t1, t2;
    t1 = t2 = recipf_(A);
    t1 *= (B - discrim);
    t2 *= (B + discrim);
    const float z1 = os[2] + os[5] * t1;
    const float z2 = os[2] + os[5] * t2;
    if (t1 < EPS_F || z1 < 0.f || z1 > 1.f) t1 = INFINITY;
    if (t2 < t1 && z2 > 0.f && z2 < 1.0f && t2 > EPS_F) return t2;
    return t1;
}

static float cylinder_hit_f(const float *invm, const float ray[6])      /* :189-222 */
{
    float os[6]; xfrm_ray_f(invm, ray, os);
    const float A = fmaf(os[3], os[3], os[4] * os[4]);
    const float B = -fmaf(os[0], os[3], os[1] * os[4]);
    const float C = fmaf(os[0], os[0], fmaf(os[1], os[1], -1.f));
    return quad_roots_select_f(A, B, C, os);
}

static float frustum_hit_f(const float *invm, const float ray[6], float td) /* :224-260 */
{
    float os[6]; xfrm_ray_f(invm, ray, os);
    const float fDir = (td - 1.0f) * os[5];
    const float fStart = (td - 1.0f) * os[2] + 1.0f;
    const float A = fmaf(os[3], os[3], fmaf(os[4], os[4], (-fDir) * fDir));
    const float B = -fmaf(os[0], os[3], fmaf(os[1], os[4], (-fStart) * fDir));
    const float C = fmaf(os[0], os[0], fmaf(os[1], os[1], (-fStart) * fStart));
    return quad_roots_select_f(A, B, C, os);
}

static float sphere_hit_f(const float *invm, const float ray[6])        /* :262-291 (Q1 kept) */
{
    float os[6]; xfrm_ray_f(invm, ray, os);
    const float A = os[3] * os[3] + os[4] * os[4] + os[5] * os[5];
    const float B = os[3] * os[0] + os[4] * os[1] + os[5] * os[2];
    const float C = os[0] * os[0] + os[1] * os[1] + os[2] * os[2] - 1.0f;
    float discrim = B * B - A * C;
    discrim = (discrim < 0.0f) ? INFINITY : discrim;    /* select(.., isless(d,0)) */
    if (discrim > 0.f) {
        float t1 = (-B - discrim) * recipf_(A);
        float t2 = (-B + discrim) * recipf_(A);
        if (t2 < t1 && t2 > EPS_F) return t2;
        else if (t1 > EPS_F) return t1;
    }
    return INFINITY;
}

float rfo_hit_f32(int32_t type, const float invm[12], float util1, float util2, const float ray[6])
{
    (void)util2;
    switch (type) {                                     /* findObject :452-461 */
    case RFO_RECTANGLE: return rectangle_hit_f(invm, ray);
    case RFO_DISC:      return disc_hit_f(invm, ray);
    case RFO_ANNULUS:   return annulus_hit_f(invm, ray, util1, util2);
    case RFO_CYLINDER:  return cylinder_hit_f(invm, ray);
    case RFO_SPHERE:    return sphere_hit_f(invm, ray);
    case RFO_FRUSTUM:   return frustum_hit_f(invm, ray, util1);
    default:            return INFINITY;
    }
}

/* ------------------------------------------------------------------ */
/* COMPAT samplers :298-443                                            */
/* ------------------------------------------------------------------ */
static void finish_ray_f(const rfo_prim *p, const float r4[4], const float osn[3], float ray[6])
{
    float o[3], wn[3];
    os_ray_direction_f(r4, o);
    transform_normal_f(p->invm, osn, wn);
    align_ray_direction_f(ray + 3, o, wn);
}

void rfo_emit_f32(const rfo_prim *p, const float r4[4], float ray[6])
{
    float v[3], n[3], o[3], wn[3], sn, cs, t;
    switch (p->type) {
    case RFO_DISC:                                                  /* :298-321 */
        t = r4[0] * M_2PI_F;
        rfo_sincosf(t, &sn, &cs); v[0] = sn; v[1] = cs;
        t = sqrtf(r4[1]);
        v[0] = t * v[0]; v[1] = t * v[1]; v[2] = 0.f;
        transform_point_f(p->m, v, ray);
        n[0] = n[1] = 0.f; n[2] = 1.f;
        finish_ray_f(p, r4, n, ray);
        break;
    case RFO_ANNULUS:                                               /* :323-345 */
        rfo_sincosf(r4[0] * M_2PI_F, &sn, &cs); v[0] = sn; v[1] = cs;
        t = (p->util2 - p->util1) * r4[1] + p->util1;
        t = sqrtf(t);
        v[0] = t * v[0]; v[1] = t * v[1]; v[2] = 0.f;
        transform_point_f(p->m, v, ray);
        n[0] = n[1] = 0.f; n[2] = 1.f;
        finish_ray_f(p, r4, n, ray);
        break;
    case RFO_CYLINDER:                                              /* :347-365 */
        rfo_sincosf(r4[0] * M_2PI_F, &sn, &cs); v[1] = sn; v[0] = cs;
        v[2] = r4[1];
        transform_point_f(p->m, v, ray);
        n[0] = v[0] * p->util2; n[1] = v[1] * p->util2; n[2] = 0.f;
        transform_normal_f(p->invm, n, wn);
        os_ray_direction_f(r4, o);
        align_ray_direction_f(ray + 3, o, wn);
        break;
    case RFO_FRUSTUM: {                                             /* :369-393 */
        const float u1 = p->util1;
        v[2] = (sqrtf((1.0f - r4[0]) + (r4[0] * u1) * u1) - 1.0f) * recipf_(u1 - 1.0f);
        t = 1.0f + v[2] * (u1 - 1.0f);
        rfo_sincosf(M_2PI_F * r4[1], &sn, &cs); v[1] = sn; v[0] = cs;
        v[0] *= t; v[1] *= t;
        transform_point_f(p->m, v, ray);
        n[0] = v[0] * p->util2; n[1] = v[1] * p->util2;
        n[2] = ((1.f - u1) * (1.f + (u1 - 1.f) * v[2])) * p->util2;
        transform_normal_f(p->invm, n, wn);
        os_ray_direction_f(r4, o);
        align_ray_direction_f(ray + 3, o, wn);
        break;
    }
    case RFO_RECTANGLE:                                             /* :399-417 */
        v[0] = 2.0f * r4[0] - 1.0f;
        v[1] = 2.0f * r4[1] - 1.0f;
        v[2] = 0.f;
        transform_point_f(p->m, v, ray);
        n[0] = n[1] = 0.f; n[2] = 1.f;
        finish_ray_f(p, r4, n, ray);
        break;
    case RFO_SPHERE:                                                /* :419-443 */
        v[2] = 2.0f * r4[0] - 1.0f;
        t = sqrtf(1.0f - v[2] * v[2]);
        rfo_sincosf(M_2PI_F * r4[1], &sn, &cs); v[1] = sn; v[0] = cs;
        v[0] *= t; v[1] *= t;
        transform_point_f(p->m, v, ray);
        n[0] = v[0] * p->util2; n[1] = v[1] * p->util2; n[2] = v[2] * p->util2;
        transform_normal_f(p->invm, n, wn);
        os_ray_direction_f(r4, o);
        align_ray_direction_f(ray + 3, o, wn);
        break;
    default:
        memset(ray, 0, 6 * sizeof(float));
        break;
    }
}

/* findObject :445-472 (strict '<': lowest index wins exact ties). */
int32_t rfo_find_f32(const rfo_prim *prims, int32_t start, int32_t end, const float ray[6])
{
    int32_t obj = -1;
    float tbest = INFINITY;
    for (int32_t w = start; w < end; ++w) {
        const rfo_prim *q = prims + w;
        float t = rfo_hit_f32(q->type, q->invm, q->util1, q->util2, ray);
        if (t < tbest) { tbest = t; obj = w; }
    }
    return obj;
}

/* runBVGPU per-ray body :578-624 (stale objectIndex quirk kept). */
int32_t rfo_find_bv_f32(const rfo_prim *prims, const rfo_bv *bvs, int32_t n_bvs, const float ray[6])
{
    int32_t bvIndex = -1, objectIndex = -1;
    float tbest = INFINITY;
    for (int32_t w = 0; w < n_bvs; ++w) {
        const rfo_bv *b = bvs + w;
        float t = rfo_hit_f32(b->type, b->invm, b->util1, b->util2, ray);
        if (t < tbest) {
            if (b->settings == 1) {
                objectIndex = rfo_find_f32(prims, b->objStartIndex, b->objEndIndex, ray);
                if (objectIndex != -1) { tbest = t; bvIndex = w; }
            } else {
                tbest = t; bvIndex = w;
            }
        }
    }
    if (bvIndex != -1 && objectIndex == -1)
        objectIndex = rfo_find_f32(prims, bvs[bvIndex].objStartIndex, bvs[bvIndex].objEndIndex, ray);
    return objectIndex;
}

/* ------------------------------------------------------------------ */
/* FP64 physics semantics                                              */
/* ------------------------------------------------------------------ */
static void xfrm_ray_d(const float *invm, const double ray[6], double os[6])
{
    for (int r = 0; r < 3; ++r) {
        const double m0 = invm[4 * r], m1 = invm[4 * r + 1], m2 = invm[4 * r + 2], m3 = invm[4 * r + 3];
        os[r] = ((m0 * ray[0] + m1 * ray[1]) + m2 * ray[2]) + m3;
        os[3 + r] = (m0 * ray[3] + m1 * ray[4]) + m2 * ray[5];
    }
}

/* Smallest root t > EPS of A t^2 + 2 B t + C = 0 whose object-space z lies
 * in [0,1] (when zcheck), else +INF.  Stable root pair q/A, C/q. */
static double quad_select_d(double A, double B, double C, const double os[6], int zcheck)
{
    double r1, r2;
    if (A == 0.0) {
        if (B == 0.0) return INFINITY;
        r1 = r2 = -C / (2.0 * B);
    } else {
        const double D = B * B - A * C;
        if (D < 0.0) return INFINITY;
        const double sq = sqrt(D);
        const double q = -(B + (B >= 0.0 ? sq : -sq));
        if (q == 0.0) return INFINITY;
        r1 = q / A;
        r2 = C / q;
    }
    const double lo = r1 < r2 ? r1 : r2;
    const double hi = r1 < r2 ? r2 : r1;
    if (lo > EPS_D) {
        const double z = os[2] + os[5] * lo;
        if (!zcheck || (z >= 0.0 && z <= 1.0)) return lo;
    }
    if (hi > EPS_D) {
        const double z = os[2] + os[5] * hi;
        if (!zcheck || (z >= 0.0 && z <= 1.0)) return hi;
    }
    return INFINITY;
}

double rfo_hit_f64(int32_t type, const float invm[12], float util1, float util2, const double ray[6])
{
    double os[6];
    xfrm_ray_d(invm, ray, os);
    switch (type) {
    case RFO_RECTANGLE: case RFO_DISC: case RFO_ANNULUS: {
        if (os[5] == 0.0) return INFINITY;
        const double t = -os[2] / os[5];
        if (!(t > EPS_D)) return INFINITY;
        const double hx = os[0] + os[3] * t;
        const double hy = os[1] + os[4] * t;
        if (type == RFO_RECTANGLE) return (fabs(hx) <= 1.0 && fabs(hy) <= 1.0) ? t : INFINITY;
        const double e = hx * hx + hy * hy;
        if (type == RFO_DISC) return e <= 1.0 ? t : INFINITY;
        return (e >= (double)util1 && e <= (double)util2) ? t : INFINITY;
    }
    case RFO_CYLINDER: {
        const double A = os[3] * os[3] + os[4] * os[4];
        const double B = os[0] * os[3] + os[1] * os[4];
        const double C = (os[0] * os[0] + os[1] * os[1]) - 1.0;
        return quad_select_d(A, B, C, os, 1);
    }
    case RFO_FRUSTUM: {
        const double k = (double)util1 - 1.0;
        const double ws = 1.0 + k * os[2];
        const double wc = k * os[5];
        const double A = (os[3] * os[3] + os[4] * os[4]) - wc * wc;
        const double B = (os[0] * os[3] + os[1] * os[4]) - ws * wc;
        const double C = (os[0] * os[0] + os[1] * os[1]) - ws * ws;
        return quad_select_d(A, B, C, os, 1);
    }
    case RFO_SPHERE: {
        const double A = (os[3] * os[3] + os[4] * os[4]) + os[5] * os[5];
        const double B = (os[0] * os[3] + os[1] * os[4]) + os[2] * os[5];
        const double C = ((os[0] * os[0] + os[1] * os[1]) + os[2] * os[2]) - 1.0;
        return quad_select_d(A, B, C, os, 0);
    }
    default:
        return INFINITY;
    }
}

/* Emission frame: n = normalize(invm^T n_os); tangent = the world image of a
 * fixed object axis, made orthogonal to n (Gram-Schmidt); c = o.x t1 + o.y t2
 * + o.z n with t2 = n x t1.  Object axis: x for planar types, the axis (z)
 * for cylinder/frustum, the axis least aligned with n_os for spheres. */
static void emit_frame_d(const rfo_prim *p, const double osn[3], const double o[3], double c[3])
{
    const float *iv = p->invm, *m = p->m;
    double n0 = (iv[0] * osn[0] + iv[4] * osn[1]) + iv[8] * osn[2];
    double n1 = (iv[1] * osn[0] + iv[5] * osn[1]) + iv[9] * osn[2];
    double n2 = (iv[2] * osn[0] + iv[6] * osn[1]) + iv[10] * osn[2];
    double inv = 1.0 / sqrt((n0 * n0 + n1 * n1) + n2 * n2);
    n0 *= inv; n1 *= inv; n2 *= inv;
    int k;
    if (p->type == RFO_CYLINDER || p->type == RFO_FRUSTUM) k = 2;
    else if (p->type == RFO_SPHERE) {
        const double a0 = fabs(osn[0]), a1 = fabs(osn[1]), a2 = fabs(osn[2]);
        k = (a0 <= a1 && a0 <= a2) ? 0 : (a1 <= a2 ? 1 : 2);
    } else k = 0;
    const double a0 = m[k], a1 = m[4 + k], a2 = m[8 + k];
    const double an = (a0 * n0 + a1 * n1) + a2 * n2;
    double t0 = a0 - an * n0, t1 = a1 - an * n1, t2 = a2 - an * n2;
    inv = 1.0 / sqrt((t0 * t0 + t1 * t1) + t2 * t2);
    t0 *= inv; t1 *= inv; t2 *= inv;
    const double b0 = n1 * t2 - n2 * t1;
    const double b1 = n2 * t0 - n0 * t2;
    const double b2 = n0 * t1 - n1 * t0;
    c[0] = (o[0] * t0 + o[1] * b0) + o[2] * n0;
    c[1] = (o[0] * t1 + o[1] * b1) + o[2] * n1;
    c[2] = (o[0] * t2 + o[1] * b2) + o[2] * n2;
}

void rfo_emit_f64(const rfo_prim *p, const double r4[4], double ray[6])
{
    double v[3], n[3], o[3], sn, cs, t;
    /* cosine-weighted local direction (osRayDirection :111-121) */
    rfo_sincos64(r4[2] * TWO_PI_D, &sn, &cs);
    t = sqrt(r4[3]);
    o[0] = t * cs; o[1] = t * sn; o[2] = sqrt(1.0 - r4[3]);
    switch (p->type) {
    case RFO_DISC:
        rfo_sincos64(r4[0] * TWO_PI_D, &sn, &cs);
        t = sqrt(r4[1]);
        v[0] = t * sn; v[1] = t * cs; v[2] = 0.0;
        n[0] = n[1] = 0.0; n[2] = 1.0;
        break;
    case RFO_ANNULUS:
        rfo_sincos64(r4[0] * TWO_PI_D, &sn, &cs);
        t = sqrt(((double)p->util2 - (double)p->util1) * r4[1] + (double)p->util1);
        v[0] = t * sn; v[1] = t * cs; v[2] = 0.0;
        n[0] = n[1] = 0.0; n[2] = 1.0;
        break;
    case RFO_CYLINDER:
        rfo_sincos64(r4[0] * TWO_PI_D, &sn, &cs);
        v[0] = cs; v[1] = sn; v[2] = r4[1];
        n[0] = cs * p->util2; n[1] = sn * p->util2; n[2] = 0.0;
        break;
    case RFO_FRUSTUM: {
        const double u1 = p->util1;
        const double a = r4[0] * (u1 * u1 - 1.0);
        v[2] = (r4[0] * (u1 + 1.0)) / (sqrt(1.0 + a) + 1.0);
        const double rho = 1.0 + v[2] * (u1 - 1.0);
        rfo_sincos64(r4[1] * TWO_PI_D, &sn, &cs);
        v[0] = rho * cs; v[1] = rho * sn;
        n[0] = v[0] * p->util2; n[1] = v[1] * p->util2; n[2] = ((1.0 - u1) * rho) * p->util2;
        break;
    }
    case RFO_RECTANGLE:
        v[0] = 2.0 * r4[0] - 1.0; v[1] = 2.0 * r4[1] - 1.0; v[2] = 0.0;
        n[0] = n[1] = 0.0; n[2] = 1.0;
        break;
    case RFO_SPHERE: {
        v[2] = 2.0 * r4[0] - 1.0;
        const double rho = sqrt((1.0 - v[2]) * (1.0 + v[2]));
        rfo_sincos64(r4[1] * TWO_PI_D, &sn, &cs);
        v[0] = rho * cs; v[1] = rho * sn;
        n[0] = v[0] * p->util2; n[1] = v[1] * p->util2; n[2] = v[2] * p->util2;
        break;
    }
    default:
        memset(ray, 0, 6 * sizeof(double));
        return;
    }
    for (int r = 0; r < 3; ++r) {
        const float *row = p->m + 4 * r;
        ray[r] = (((double)row[0] * v[0] + (double)row[1] * v[1]) + (double)row[2] * v[2]) + (double)row[3];
    }
    emit_frame_d(p, n, o, ray + 3);
}

int32_t rfo_find_f64(const rfo_prim *prims, int32_t start, int32_t end, const double ray[6])
{
    int32_t obj = -1;
    double tbest = INFINITY;
    for (int32_t w = start; w < end; ++w) {
        const rfo_prim *q = prims + w;
        double t = rfo_hit_f64(q->type, q->invm, q->util1, q->util2, ray);
        if (t < tbest) { tbest = t; obj = w; }
    }
    return obj;
}

int32_t rfo_find_bv_f64(const rfo_prim *prims, const rfo_bv *bvs, int32_t n_bvs, const double ray[6])
{
    int32_t bvIndex = -1, objectIndex = -1;
    double tbest = INFINITY;
    for (int32_t w = 0; w < n_bvs; ++w) {
        const rfo_bv *b = bvs + w;
        double t = rfo_hit_f64(b->type, b->invm, b->util1, b->util2, ray);
        if (t < tbest) {
            if (b->settings == 1) {
                objectIndex = rfo_find_f64(prims, b->objStartIndex, b->objEndIndex, ray);
                if (objectIndex != -1) { tbest = t; bvIndex = w; }
            } else {
                tbest = t; bvIndex = w;
            }
        }
    }
    if (bvIndex != -1 && objectIndex == -1)
        objectIndex = rfo_find_f64(prims, bvs[bvIndex].objStartIndex, bvs[bvIndex].objEndIndex, ray);
    return objectIndex;
}

/* ------------------------------------------------------------------ */
/* Launch geometry: main.cpp:101,172-173                               */
/* ------------------------------------------------------------------ */
int64_t rfo_nthreads(int32_t ray_density, int32_t n_samples, int32_t count, int32_t wgsize)
{
    const int64_t prod = (int64_t)ray_density * (int64_t)n_samples;
    if (prod > INT32_MAX || prod < INT32_MIN || count <= 0 || wgsize <= 0) return -1;
    const int64_t nt = prod / count;          /* int arithmetic, truncating */
    if (nt < 0) return -1;                    /* size_t wrap in the reference */
    const int64_t rem = nt % wgsize;
    return rem > wgsize / 2 ? nt + wgsize - rem : nt - rem;
}

/* ------------------------------------------------------------------ */
/* Work-item driver: runGPU :481-527 / runBVGPU :530-626               */
/* ------------------------------------------------------------------ */
static int32_t trace_one(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
                         const rfo_opts *o, const float rf[4], const double rd[4],
                         const rfo_prim *em)
{
    if (o->mode == RFO_MODE_COMPAT) {
        float ray[6];
        rfo_emit_f32(em, rf, ray);
        return o->use_bvs ? rfo_find_bv_f32(prims, bvs, n_bvs, ray) : rfo_find_f32(prims, 0, n, ray);
    } else {
        double ray[6];
        rfo_emit_f64(em, rd, ray);
        return o->use_bvs ? rfo_find_bv_f64(prims, bvs, n_bvs, ray) : rfo_find_f64(prims, 0, n, ray);
    }
}

/* The 4 rays of a work item use cyclic rotations of one Philox block,
 * rotated before every ray (rands.yzwx, :502-506). */
static void workitem_rays(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
                          const rfo_opts *o, int32_t emitter, int64_t wi, int32_t out[4])
{
    uint32_t u[4];
    rfo_workitem_u32(o->rng, o->seed, o->geometry, emitter, wi, u);
    float f[4];
    for (int i = 0; i < 4; ++i) f[i] = rfo_u01_24(u[i]);
    const rfo_prim *em = prims + emitter;
    for (int v = 0; v < 4; ++v) {
        float rf[4]; double rd[4];
        for (int i = 0; i < 4; ++i) { rf[i] = f[(i + v + 1) & 3]; rd[i] = (double)rf[i]; }
        out[v] = trace_one(prims, n, bvs, n_bvs, o, rf, rd, em);
    }
}

void rfo_trace(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
               const rfo_opts *opts, int32_t emitter, int64_t wi0, int64_t n_wi,
               int32_t *idx_out)
{
    for (int64_t i = 0; i < n_wi; ++i)
        workitem_rays(prims, n, bvs, n_bvs, opts, emitter, wi0 + i, idx_out + 4 * i);
}

static int validate(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs)
{
    if (n <= 0 || !prims) return -1;
    for (int32_t i = 0; i < n; ++i)
        if (prims[i].type < 0 || prims[i].type >= RFO_NONE) return -2;
    for (int32_t i = 0; i < n_bvs; ++i) {
        if (bvs[i].type < 0 || bvs[i].type >= RFO_NONE) return -2;
        if (bvs[i].objStartIndex < 0 || bvs[i].objEndIndex > n ||
            bvs[i].objStartIndex > bvs[i].objEndIndex) return -3;
    }
    return 0;
}

int32_t rfo_solve_range(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
                        const rfo_opts *opts, int32_t emitter, int64_t wi0, int64_t n_wi,
                        int64_t *hits_row)
{
    int rc = validate(prims, n, bvs, n_bvs);
    if (rc) return rc;
    for (int64_t i = 0; i < n_wi; ++i) {
        int32_t idx[4];
        workitem_rays(prims, n, bvs, n_bvs, opts, emitter, wi0 + i, idx);
        for (int v = 0; v < 4; ++v)
            if (idx[v] >= 0) hits_row[idx[v]]++;
    }
    return 0;
}

/* ---- threaded solve ---- */
typedef struct {
    const rfo_prim *prims; int32_t n; const rfo_bv *bvs; int32_t n_bvs;
    const rfo_opts *opts; int32_t e0, e1; const int64_t *n_wi; int64_t *hits;
    int64_t block; int64_t n_tasks; int64_t *task_prefix;   /* per-emitter task offsets */
    volatile int64_t next;
} solve_ctx;

static void *solve_worker(void *arg)
{
    solve_ctx *S = (solve_ctx *)arg;
    int64_t *row = (int64_t *)calloc((size_t)S->n, sizeof(int64_t));
    int32_t cur_e = -1;
    for (;;) {
        int64_t task = __atomic_fetch_add(&S->next, 1, __ATOMIC_RELAXED);
        if (task >= S->n_tasks) break;
        int32_t e = S->e0;
        while (S->task_prefix[e - S->e0 + 1] <= task) ++e;
        if (e != cur_e && cur_e >= 0) {
            for (int32_t j = 0; j < S->n; ++j)
                if (row[j]) { __atomic_fetch_add(&S->hits[(int64_t)cur_e * S->n + j], row[j], __ATOMIC_RELAXED); row[j] = 0; }
        }
        cur_e = e;
        const int64_t local = task - S->task_prefix[e - S->e0];
        const int64_t wi0 = local * S->block;
        int64_t cnt = S->n_wi[e] - wi0;
        if (cnt > S->block) cnt = S->block;
        for (int64_t i = 0; i < cnt; ++i) {
            int32_t idx[4];
            workitem_rays(S->prims, S->n, S->bvs, S->n_bvs, S->opts, e, wi0 + i, idx);
            for (int v = 0; v < 4; ++v)
                if (idx[v] >= 0) row[idx[v]]++;
        }
    }
    if (cur_e >= 0)
        for (int32_t j = 0; j < S->n; ++j)
            if (row[j]) __atomic_fetch_add(&S->hits[(int64_t)cur_e * S->n + j], row[j], __ATOMIC_RELAXED);
    free(row);
    return NULL;
}

int32_t rfo_solve(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
                  const rfo_opts *opts, int32_t e0, int32_t e1, const int64_t *n_wi,
                  int64_t *hits, int64_t *rays)
{
    int rc = validate(prims, n, bvs, n_bvs);
    if (rc) return rc;
    if (e0 < 0 || e1 > n || e0 > e1) return -4;
    solve_ctx S;
    memset(&S, 0, sizeof S);
    S.prims = prims; S.n = n; S.bvs = bvs; S.n_bvs = n_bvs; S.opts = opts;
    S.e0 = e0; S.e1 = e1; S.n_wi = n_wi; S.hits = hits;
    S.block = 1024;
    S.task_prefix = (int64_t *)calloc((size_t)(e1 - e0 + 1), sizeof(int64_t));
    for (int32_t e = e0; e < e1; ++e) {
        if (n_wi[e] < 0) { free(S.task_prefix); return -5; }
        S.task_prefix[e - e0 + 1] = S.task_prefix[e - e0] + (n_wi[e] + S.block - 1) / S.block;
        if (rays) rays[e] = 4 * n_wi[e];
    }
    S.n_tasks = S.task_prefix[e1 - e0];
    int nt = opts->n_threads > 0 ? opts->n_threads : 1;
    if (nt > 256) nt = 256;
    pthread_t th[256];
    for (int i = 0; i < nt; ++i) pthread_create(&th[i], NULL, solve_worker, &S);
    for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
    free(S.task_prefix);
    return 0;
}
