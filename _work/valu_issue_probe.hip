// VALU issue-rate probe for gfx950: FP32 instruction streams of different
// mixes at 2-8 waves per SIMD.  Every thread writes one float (no other
// memory traffic); loops have fixed trip counts.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>

#define CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e_), __LINE__); return 1; } } while (0)

template <int MODE>
__global__ void __launch_bounds__(256) k(float *out, float a, float b, int iters, unsigned long long *clk)
{
    float v[8];
    for (int i = 0; i < 8; ++i) v[i] = threadIdx.x * 1e-3f + i;
    float acc = 0.f;
    unsigned long long t0 = __builtin_amdgcn_s_memtime();
    unsigned long long r0 = __builtin_amdgcn_s_memrealtime();
#pragma unroll 4
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (MODE == 0) {            // FMA, VGPR operands only
                v[i] = __builtin_fmaf(v[i], v[(i + 1) & 7], v[(i + 3) & 7]);
            } else if constexpr (MODE == 1) {     // FMA with SGPR operands
                v[i] = __builtin_fmaf(v[i], a, b);
            } else if constexpr (MODE == 2) {     // mul
                v[i] = v[i] * v[(i + 1) & 7];
            } else if constexpr (MODE == 3) {     // FMA + compare/select pair
                const float t = __builtin_fmaf(v[i], a, b);
                v[i] = (t < v[(i + 1) & 7]) ? t : v[(i + 2) & 7];
            } else if constexpr (MODE == 4) {     // FMA + rcp every 8th
                v[i] = (i == 0) ? __builtin_amdgcn_rcpf(v[i]) : __builtin_fmaf(v[i], a, b);
            } else if constexpr (MODE == 5) {     // FMA + sqrt every 4th
                v[i] = (i & 3) == 0 ? __builtin_amdgcn_sqrtf(__builtin_fabsf(v[i])) : __builtin_fmaf(v[i], a, b);
            } else if constexpr (MODE == 6) {     // FMA(vgpr) + cmp to SGPR pair + cndmask_e64 on it
                unsigned long long m;
                const float t = __builtin_fmaf(v[i], v[(i + 1) & 7], v[(i + 3) & 7]);
                asm volatile("v_cmp_lt_f32_e64 %0, %1, %2" : "=s"(m) : "v"(t), "v"(v[(i + 2) & 7]));
                asm volatile("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(v[i]) : "v"(t), "v"(v[(i + 5) & 7]), "s"(m));
            } else if constexpr (MODE == 7) {     // FMA(vgpr) + cmp to VCC + cndmask_e32 on VCC
                const float t = __builtin_fmaf(v[i], v[(i + 1) & 7], v[(i + 3) & 7]);
                asm volatile("v_cmp_lt_f32_e32 vcc, %1, %2\n\tv_cndmask_b32_e32 %0, %1, %3, vcc"
                             : "=v"(v[i]) : "v"(t), "v"(v[(i + 2) & 7]), "v"(v[(i + 5) & 7]) : "vcc");
            } else if constexpr (MODE == 8) {     // FMA(vgpr) + v_mov from SGPR
                float c;
                asm volatile("v_mov_b32 %0, %1" : "=v"(c) : "s"(a));
                v[i] = __builtin_fmaf(v[i], v[(i + 1) & 7], c);
            } else if constexpr (MODE == 9) {     // FMA(vgpr) + sqrt(vgpr) every 8th
                v[i] = (i == 0) ? __builtin_amdgcn_sqrtf(__builtin_fabsf(v[i])) : __builtin_fmaf(v[i], v[(i + 1) & 7], v[(i + 3) & 7]);
            } else if constexpr (MODE == 10) {    // FMA(vgpr) + literal constant operand
                v[i] = __builtin_fmaf(v[i], 1.0001f, v[(i + 3) & 7]);
            }
        }
    }
    unsigned long long t1 = __builtin_amdgcn_s_memtime();
    unsigned long long r1 = __builtin_amdgcn_s_memrealtime();
    for (int i = 0; i < 8; ++i) acc += v[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
    if (threadIdx.x == 0 && blockIdx.x == 0) { clk[0] = t1 - t0; clk[1] = r1 - r0; }
}

template <int MODE>
int run(const char *name, int n_cu, float *d_out, unsigned long long *d_clk, int wps)
{
    // wps waves of 64 per SIMD: groups of 256 threads (4 waves), n_cu * wps groups
    const int groups = n_cu * wps, block = 256, iters = 4096;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0)); CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k<MODE>, dim3(groups), dim3(block), 0, 0, d_out, 1.0001f, 0.5f, 16, d_clk);
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0));
    hipLaunchKernelGGL(k<MODE>, dim3(groups), dim3(block), 0, 0, d_out, 1.0001f, 0.5f, iters, d_clk);
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    float ms; CHK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long clk[2]; CHK(hipMemcpy(clk, d_clk, sizeof clk, hipMemcpyDeviceToHost));
    const double ghz = clk[1] ? (double)clk[0] / (clk[1] / 100.0) / 1e3 : 0.0;   // memrealtime at 100 MHz
    const double instr = (double)groups * block / 64 * iters * 8 * (MODE == 3 ? 3 : (MODE == 6 || MODE == 7) ? 3 : MODE == 8 ? 3 : 1);   // wave VALU instructions
    const double per_simd_cycle = instr / (n_cu * 4.0) / (ms * 1e-3 * ghz * 1e9);
    printf("%-22s waves/SIMD %d: %8.3f ms, %.2f GHz, %.3f wave-instr per SIMD-cycle (lane-ops/clk/SIMD %.1f)\n",
           name, wps, ms, ghz, per_simd_cycle, per_simd_cycle * 64);
    CHK(hipEventDestroy(e0)); CHK(hipEventDestroy(e1));
    return 0;
}

int main()
{
    int dev = 0; hipDeviceProp_t p; if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 1;
    const int n_cu = p.multiProcessorCount;
    printf("%s, %d CUs\n", p.gcnArchName, n_cu);
    float *d_out; unsigned long long *d_clk;
    CHK(hipMalloc(&d_out, (size_t)n_cu * 8 * 256 * sizeof(float)));
    CHK(hipMalloc(&d_clk, 2 * sizeof(unsigned long long)));
    for (int wps : {2, 4, 6, 8}) {
        if (run<0>("fma vgpr", n_cu, d_out, d_clk, wps)) return 1;
        if (run<1>("fma sgpr", n_cu, d_out, d_clk, wps)) return 1;
        if (run<2>("mul", n_cu, d_out, d_clk, wps)) return 1;
        if (run<3>("fma+cmp+cndmask", n_cu, d_out, d_clk, wps)) return 1;
        if (run<4>("fma+rcp(1/8)", n_cu, d_out, d_clk, wps)) return 1;
        if (run<5>("fma+sqrt(1/4)", n_cu, d_out, d_clk, wps)) return 1;
        if (run<6>("fma+cmp_e64+cnd_e64", n_cu, d_out, d_clk, wps)) return 1;
        if (run<7>("fma+cmp_vcc+cnd_vcc", n_cu, d_out, d_clk, wps)) return 1;
        if (run<8>("fma+v_mov_from_sgpr", n_cu, d_out, d_clk, wps)) return 1;
        if (run<9>("fma vgpr+sqrt(1/8)", n_cu, d_out, d_clk, wps)) return 1;
        if (run<10>("fma literal", n_cu, d_out, d_clk, wps)) return 1;
    }
    CHK(hipFree(d_out)); CHK(hipFree(d_clk));
    return 0;
}
