// VALU issue-rate probe for gfx950: every lane runs 8 independent dependency chains of
// v_fma_f32, of v_exp_f32, or of both alternating, for a fixed number of iterations, over
// 4096 workgroups of 256 lanes. Prints lane-operations per second for each mix. Each lane
// writes one float of a buffer sized for every lane, so all accesses stay in bounds.
#include <hip/hip_runtime.h>
#include <cstdio>

template <int KIND>
__global__ __launch_bounds__(256) void chains(float* out, int iters, float a) {
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = (threadIdx.x + i) * 1e-3f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (KIND != 1) x[i] = __builtin_fmaf(x[i], a, 0.5f);
            if (KIND != 0) x[i] = __builtin_amdgcn_exp2f(-x[i]);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += x[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

int main() {
    const int blocks = 256 * 16, threads = 256, iters = 4096;
    float* d = nullptr;
    CHECK(hipMalloc(&d, sizeof(float) * blocks * threads));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const char* names[3] = {"v_fma_f32", "v_exp_f32", "v_fma_f32 + v_exp_f32"};
    for (int kind = 0; kind < 3; ++kind) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
            CHECK(hipEventRecord(e0));
            if (kind == 0) hipLaunchKernelGGL(chains<0>, dim3(blocks), dim3(threads), 0, 0, d, iters, 0.999f);
            if (kind == 1) hipLaunchKernelGGL(chains<1>, dim3(blocks), dim3(threads), 0, 0, d, iters, 0.999f);
            if (kind == 2) hipLaunchKernelGGL(chains<2>, dim3(blocks), dim3(threads), 0, 0, d, iters, 0.999f);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        const double ops = (double)blocks * threads * iters * 8 * (kind == 2 ? 2 : 1);
        printf("%-22s %8.3f ms  %6.2f T lane-ops/s\n", names[kind], ms, ops / ms / 1e9);
    }
    CHECK(hipFree(d));
    return 0;
}
