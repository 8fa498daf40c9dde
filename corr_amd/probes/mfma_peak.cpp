// MFMA issue-rate probe for gfx950: every wave runs ITER x (independent MFMA
// chains) from registers, so the measured rate is the matrix-core ceiling that
// the dense kernels of this library are held against.  Prints one line per
// instruction: TFLOP/s over 2048 workgroups of 4 waves.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
constexpr int ITER = 4096;

__global__ __launch_bounds__(256) void k_bf16_16(float* out) {
    bf8 a, b;
    for (int i = 0; i < 8; ++i) {
        a[i] = (__bf16)(threadIdx.x * 0.001f + i);
        b[i] = (__bf16)(i * 0.5f);
    }
    f4 acc[8];
    for (int j = 0; j < 8; ++j) acc[j] = f4{0, 0, 0, 0};
    for (int it = 0; it < ITER; ++it)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_bf16_32(float* out) {
    bf8 a, b;
    for (int i = 0; i < 8; ++i) {
        a[i] = (__bf16)(threadIdx.x * 0.001f + i);
        b[i] = (__bf16)(i * 0.5f);
    }
    f16v acc[4];
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 16; ++r) acc[j][r] = 0;
    for (int it = 0; it < ITER; ++it)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
    float s = 0;
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 16; ++r) s += acc[j][r];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_f64(float* out) {
    double a = threadIdx.x * 0.001, b = 0.5;
    d4 acc[8];
    for (int j = 0; j < 8; ++j) acc[j] = d4{0, 0, 0, 0};
    for (int it = 0; it < ITER; ++it)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    double s = 0;
    for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * 256 + threadIdx.x] = (float)s;
}

__global__ __launch_bounds__(256) void k_f32(float* out) {
    float a = threadIdx.x * 0.001f, b = 0.5f;
    f4 acc[8];
    for (int j = 0; j < 8; ++j) acc[j] = f4{0, 0, 0, 0};
    for (int it = 0; it < ITER; ++it)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int main() {
    const int blocks = 256 * 8;   // 8 workgroups of 4 waves per CU
    float* out = nullptr;
    if (hipMalloc(&out, blocks * 256 * sizeof(float)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    struct K {
        const char* name;
        void (*fn)(float*);
        double flops_per_inst;
        int per_iter;
        double spec;
    };
    const K ks[] = {{"v_mfma_f32_16x16x32_bf16", k_bf16_16, 16.0 * 16 * 32 * 2, 8, 2500.0},
                    {"v_mfma_f32_32x32x16_bf16", k_bf16_32, 32.0 * 32 * 16 * 2, 4, 2500.0},
                    {"v_mfma_f64_16x16x4_f64", k_f64, 16.0 * 16 * 4 * 2, 8, 78.6},
                    {"v_mfma_f32_16x16x4_f32", k_f32, 16.0 * 16 * 4 * 2, 8, 157.3}};
    for (const K& k : ks) {
        hipLaunchKernelGGL(k.fn, dim3(blocks), dim3(256), 0, 0, out);
        if (hipEventRecord(e0) != hipSuccess) return 1;
        hipLaunchKernelGGL(k.fn, dim3(blocks), dim3(256), 0, 0, out);
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        const double fl = (double)blocks * 4 * ITER * k.per_iter * k.flops_per_inst;
        printf("%-26s %8.3f ms  %8.1f TFLOP/s  (%.0f%% of the %.1f published)\n", k.name, ms, fl / ms / 1e9,
               100.0 * fl / ms / 1e9 / k.spec, k.spec);
    }
    return hipFree(out) == hipSuccess ? 0 : 1;
}
