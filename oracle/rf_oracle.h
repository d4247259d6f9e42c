/*
 * rf_oracle.h -- CPU restatement of RayFactorCL's Monte Carlo view-factor
 * kernels (reference: RayFactorCL/RayFactorGPU.cl), used ONLY as a test
 * oracle and as the CPU baseline of the benchmark.  Nothing in the product
 * library links or calls this code.
 *
 * Two arithmetic modes are restated:
 *
 *   RFO_MODE_COMPAT  FP32, line-for-line semantics of RayFactorGPU.cl
 *                    (including its quirks: sphere test without the square
 *                    root, asymmetric root acceptance, strict '<' ties, the
 *                    isinf() frame branch), with the vendor-defined builtins
 *                    pinned to a fixed convention:
 *                      native_recip(x) -> 1.0f/x        (IEEE, correctly rounded)
 *                      native_sqrt(x)  -> sqrtf(x)      (IEEE, correctly rounded)
 *                      native_rsqrt(x) -> 1.0f/sqrtf(x)
 *                      mad(a,b,c)      -> fmaf(a,b,c)
 *                      sincos(x)       -> rfo_sincosf (own, below)
 *                      dot(a,b)        -> ((a0*b0 + a1*b1) + a2*b2) + a3*b3
 *                    and no other floating-point contraction.
 *
 *   RFO_MODE_FP64    FP64 "physics" semantics: the same sampling
 *                    parametrisation and random-number use, a corrected
 *                    sphere test, numerically stable quadratic roots, an
 *                    explicit negative-discriminant miss, and an
 *                    orthonormal emission frame built from a fixed tangent
 *                    (Gram-Schmidt) instead of the 1/(1+n_z) rotation.
 *                    This is the error oracle for the FP32 fast path.
 *
 * Parity pins (what fixes each function independently of this file):
 *   Philox4x32-10      published Random123 known-answer vectors and the
 *                      reference header compiled from its own sources
 *   u01                closed form (exact)
 *   sincos             error bound against libm double sin/cos
 *   hit tests          analytic intersections of constructed rays
 *   samplers           surface residual, |c| = 1, n.c >= 0, distribution tests
 *   findObject / BV    brute force on tiny scenes, tie invariant
 *   whole solve        Howell-catalogue closed forms, reciprocity, closure
 *   launch geometry    integer known answers of main.cpp:172-173
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-identical to RayfactorConstants.h:38-45 (112 bytes). */
typedef struct rfo_prim {
    float m[12];
    float invm[12];
    int32_t type;
    int32_t rayDensity;
    float util1;
    float util2;
} rfo_prim;

/* Byte-identical to RayfactorConstants.h:48-56 (72 bytes). */
typedef struct rfo_bv {
    float invm[12];
    int32_t type;
    int32_t settings;
    int32_t objStartIndex;
    int32_t objEndIndex;
    float util1;
    float util2;
} rfo_bv;

enum { RFO_CYLINDER = 0, RFO_FRUSTUM = 1, RFO_SPHERE = 2, RFO_ANNULUS = 3,
       RFO_RECTANGLE = 4, RFO_DISC = 5, RFO_NONE = 6 };

enum { RFO_MODE_COMPAT = 0, RFO_MODE_FP64 = 1 };
enum { RFO_RNG_REFERENCE = 0, RFO_RNG_SEEDED = 1 };

typedef struct rfo_opts {
    int32_t mode;          /* RFO_MODE_* */
    int32_t use_bvs;       /* 0: runGPU semantics, 1: runBVGPU semantics */
    int32_t rng;           /* RFO_RNG_* */
    uint32_t seed;         /* seeded RNG only */
    uint32_t geometry;     /* seeded RNG only: geometry id in a sweep */
    int32_t n_threads;     /* host threads (<=0: 1) */
} rfo_opts;

/* ---- building blocks (exported for unit tests) ---- */
void rfo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float rfo_u01_24(uint32_t u);
void rfo_sincosf(float x, float *s, float *c);
void rfo_sincos64(double x, double *s, double *c);

/* Random 4-vector for one work item (before the per-ray rotation). */
void rfo_workitem_u32(int32_t rng, uint32_t seed, uint32_t geometry, int32_t emitter,
                      int64_t wi, uint32_t out[4]);

/* Emission: ray = {s.x,s.y,s.z,c.x,c.y,c.z} in world space. */
void rfo_emit_f32(const rfo_prim *p, const float rands[4], float ray[6]);
void rfo_emit_f64(const rfo_prim *p, const double rands[4], double ray[6]);

/* One ray-primitive test; returns +INF on a miss (or NaN where the
 * reference formula yields NaN, which never wins a comparison). */
float rfo_hit_f32(int32_t type, const float invm[12], float util1, float util2, const float ray[6]);
double rfo_hit_f64(int32_t type, const float invm[12], float util1, float util2, const double ray[6]);

/* Nearest hit over [start,end) (findObject); -1 on a miss. */
int32_t rfo_find_f32(const rfo_prim *prims, int32_t start, int32_t end, const float ray[6]);
int32_t rfo_find_f64(const rfo_prim *prims, int32_t start, int32_t end, const double ray[6]);

/* BV kernel body for one ray (runBVGPU lines :578-624); -1 on a miss. */
int32_t rfo_find_bv_f32(const rfo_prim *prims, const rfo_bv *bvs, int32_t n_bvs, const float ray[6]);
int32_t rfo_find_bv_f64(const rfo_prim *prims, const rfo_bv *bvs, int32_t n_bvs, const double ray[6]);

/* Launch geometry of main.cpp:172-173: work items for one emitter.
 * Returns -1 if rayDensity*n_samples overflows int32 (undefined behaviour in
 * the reference). */
int64_t rfo_nthreads(int32_t ray_density, int32_t n_samples, int32_t count, int32_t wgsize);

/* Per-ray trace: hit index of each of the 4*n_wi rays of work items
 * [wi0, wi0+n_wi) of one emitter. */
void rfo_trace(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
               const rfo_opts *opts, int32_t emitter, int64_t wi0, int64_t n_wi,
               int32_t *idx_out);

/* Whole solve: for every emitter e in [e0,e1), n_wi[e] work items of 4 rays
 * each; hits is the caller's n*n int64 matrix (accumulated into), rays[e]
 * receives 4*n_wi[e].  Returns 0, or a negative code on invalid input. */
int32_t rfo_solve(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
                  const rfo_opts *opts, int32_t e0, int32_t e1, const int64_t *n_wi,
                  int64_t *hits, int64_t *rays);

/* Work-item range of one emitter (for chunked / sampled runs). */
int32_t rfo_solve_range(const rfo_prim *prims, int32_t n, const rfo_bv *bvs, int32_t n_bvs,
                        const rfo_opts *opts, int32_t emitter, int64_t wi0, int64_t n_wi,
                        int64_t *hits_row);

#ifdef __cplusplus
}
#endif
#endif
