// rayfactor -- drop-in replacement of RayFactorCL's main() (main.cpp:275-284):
// reads a scene XML (default "input.xml" in the working directory, as the
// reference), runs the view-factor solve on HIP device 0 and writes
// "output.txt" in the reference format.
//
//   rayfactor [scene.xml] [--math compat|fp64|fast] [--rays N] [--out PATH] [--report]
//
// --math compat (default) reproduces the reference's launch geometry and
// arithmetic; --rays N switches to an explicit budget of N rays per emitter.
// --report prints the closure (row sums) and reciprocity (A_i F_ij against
// A_j F_ji over pairs with both products >= 1e-3) of the FP64 view factors.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rayfactor.h"

static int fail(const char *what, int rc, const rf_ctx *ctx)
{
    std::fprintf(stderr, "rayfactor: %s: %s (%s)\n", what, rf_status_string(rc), rf_last_error(ctx));
    return 1;
}

int main(int argc, char **argv)
{
    std::string path = "input.xml", out = "output.txt";
    int math = RF_MATH_COMPAT;
    long long rays = 0;
    bool report = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--math") && i + 1 < argc) {
            const char *m = argv[++i];
            math = !std::strcmp(m, "fp64") ? RF_MATH_FP64 : (!std::strcmp(m, "fast") ? RF_MATH_FAST : RF_MATH_COMPAT);
        } else if (!std::strcmp(argv[i], "--rays") && i + 1 < argc) {
            rays = std::atoll(argv[++i]);
        } else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) {
            out = argv[++i];
        } else if (!std::strcmp(argv[i], "--report")) {
            report = true;
        } else {
            path = argv[i];
        }
    }
    rf_host_scene hs;
    int rc = rf_scene_load_xml(path.c_str(), &hs);
    if (rc) { std::fprintf(stderr, "rayfactor: failed to load \"%s\": %s\n", path.c_str(), rf_status_string(rc)); return 1; }
    rf_ctx *ctx = nullptr;
    rc = rf_ctx_create(0, &ctx);
    if (rc) return fail("device", rc, nullptr);
    rf_scene *scene = nullptr;
    rc = rf_scene_create(ctx, hs.prims, hs.n_prims, hs.bvs, hs.n_bvs, &scene);
    if (rc) return fail("scene", rc, ctx);
    rf_solve_opts o;
    rf_solve_opts_default(&o);
    o.math = math;
    o.use_bvs = hs.useBoundingVolumes;
    o.n_samples = (int)hs.globalRayDensity;   // float -> int as main.cpp:281
    if (rays > 0) { o.budget = RF_BUDGET_EXPLICIT; o.rays_per_emitter = rays; }
    const size_t n = (size_t)hs.n_prims;
    std::vector<int64_t> hits(n * n), nr(n);
    rf_timing t{};
    rc = rf_solve(ctx, scene, &o, hits.data(), nr.data(), &t);
    if (rc) return fail("solve", rc, ctx);
    std::vector<double> samples(n);
    for (size_t i = 0; i < n; ++i) samples[i] = (double)nr[i];
    rc = rf_write_output(out.c_str(), (int)n, hits.data(), samples.data());
    if (rc) return fail("output", rc, ctx);
    std::printf("Number of objects : %zu\nKernel time : %.3f ms\nRays : %lld\n", n, t.kernel_ms, (long long)t.rays);
    if (report) {
        std::vector<double> vf(n * n), area(n), closure(n);
        rf_vf_report r;
        rc = rf_normalise((int)n, hits.data(), nr.data(), vf.data());
        if (!rc) rc = rf_areas(hs.prims, (int)n, area.data());
        if (!rc) rc = rf_vf_check((int)n, vf.data(), area.data(), 1e-3, closure.data(), &r);
        if (rc) return fail("report", rc, ctx);
        std::printf("Closure : min %.8f max %.8f\n", r.min_closure, r.max_closure);
        if (r.pairs_compared > 0)
            std::printf("Reciprocity : %d pairs, max |AiFij - AjFji| %.3e, max relative %.3e (pair %d, %d)\n",
                        r.pairs_compared, r.max_recip_abs, r.max_recip_rel, r.worst_i, r.worst_j);
        else
            std::printf("Reciprocity : no pair with both products >= 1e-3\n");
    }
    rf_scene_destroy(scene);
    rf_ctx_destroy(ctx);
    rf_host_scene_free(&hs);
    return 0;
}
