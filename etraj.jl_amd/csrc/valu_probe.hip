// valu_probe.hip -- the vector-ALU throughput this device reaches with the
// instruction forms the engine compiles to (non-packed FP32, FP64), as a
// ceiling to set the integrator kernels' rates against. Packed FP32
// (v_pk_fma_f32) is disabled in every build of this project, so the published
// FP32 vector peak, which counts packed FMAs, is out of reach by construction.
//
// Each kernel runs 16 independent dependency chains per lane, every operand in
// a vector register, 8 wavefronts per SIMD, and is timed with HIP events (best
// of 3 runs). Output: one JSON object on stdout.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kChains = 16;
constexpr int kIters = 20000;

template <int OP>
__global__ __launch_bounds__(256) void k_probe32(float* out, float a0, float b0) {
    const float a = a0 + threadIdx.x * 1e-9f, b = b0 + threadIdx.x * 1e-10f;
    float x[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = threadIdx.x * 1e-3f + k;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            if (OP == 0) x[k] = __builtin_fmaf(x[k], a, b);      // v_fma_f32
            if (OP == 1) x[k] = x[k] + a;                        // v_add_f32
            if (OP == 2) x[k] = __builtin_amdgcn_rsqf(x[k] + b);  // v_add_f32 + v_rsq_f32
        }
    }
    float s = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += x[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe64(double* out, double a0, double b0) {
    const double a = a0 + threadIdx.x * 1e-12, b = b0 + threadIdx.x * 1e-13;
    double x[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = threadIdx.x * 1e-3 + k;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) x[k] = __builtin_fma(x[k], a, b);   // v_fma_f64
    }
    double s = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += x[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

int main() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        std::fprintf(stderr, "valu_probe: no HIP device\n");
        return 1;
    }
    const int blocks = cus * 8;   // 4 wavefronts per block: 8 per SIMD
    const size_t n = (size_t)blocks * 256;
    void* out = nullptr;
    if (hipMalloc(&out, n * sizeof(double)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    const char* names[] = {"fma_f32", "add_f32", "add_rsq_f32", "fma_f64"};
    const double insts_per_chain_step[] = {1, 1, 2, 1};
    const double flop_per_chain_step[] = {2, 1, 2, 2};   // rsq counted as 1
    std::printf("{\"cus\": %d, \"waves_per_simd\": 8, \"chains\": %d", cus, kChains);
    for (int op = 0; op < 4; ++op) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            if (hipEventRecord(e0, 0) != hipSuccess) return 1;
            switch (op) {
                case 0: hipLaunchKernelGGL(k_probe32<0>, dim3(blocks), dim3(256), 0, 0, (float*)out, 0.999f, 1e-3f); break;
                case 1: hipLaunchKernelGGL(k_probe32<1>, dim3(blocks), dim3(256), 0, 0, (float*)out, 1e-7f, 0.0f); break;
                case 2: hipLaunchKernelGGL(k_probe32<2>, dim3(blocks), dim3(256), 0, 0, (float*)out, 0.0f, 1.0f); break;
                default: hipLaunchKernelGGL(k_probe64, dim3(blocks), dim3(256), 0, 0, (double*)out, 0.999, 1e-3); break;
            }
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, 0) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess)
                return 1;
            float ms = 0;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
            if (ms < best) best = ms;
        }
        const double steps = (double)n * kIters * kChains;   // lane chain-steps
        const double wave_insts = steps / 64 * insts_per_chain_step[op];
        std::printf(", \"%s\": {\"ms\": %.4f, \"wave_insts_per_simd_per_ns\": %.5f, \"tflops\": %.2f}", names[op], best,
                    wave_insts / (cus * 4.0) / (best * 1e6), steps * flop_per_chain_step[op] / (best * 1e-3) / 1e12);
    }
    std::printf("}\n");
    (void)hipFree(out);
    return 0;
}
