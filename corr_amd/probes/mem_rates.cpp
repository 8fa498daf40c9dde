// LDS and L2 rate probe for gfx950.  One 512-thread workgroup per CU (the LDS
// footprint allows no second one); every address pattern is bank-conflict-free
// and in bounds.  Prints bytes per nanosecond per CU for:
//   ds_read_b128, ds_write_b128, global_load_dwordx4 and
//   global_load_lds_dwordx4 from a 2 MiB buffer that stays in L2.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

typedef unsigned u4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr;
typedef __attribute__((address_space(1))) const void* gbl_ptr;

constexpr int LDS_BYTES = 96 * 1024;   // > 80 KiB: one workgroup per CU
constexpr int ITER = 2048;
constexpr int BUF = 2 << 20;           // L2-resident source
constexpr int CHUNKS = BUF / 1024;
constexpr int PASSES = 8;

// MODE 0: 16 ds_read_b128 per iteration per wave; 1: 16 ds_write_b128.
template <int MODE>
__global__ __launch_bounds__(512) void lds_kernel(unsigned* out) {
    __shared__ __attribute__((aligned(16))) unsigned char L[LDS_BYTES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < LDS_BYTES / 4; i += 512) reinterpret_cast<unsigned*>(L)[i] = i;
    __syncthreads();
    u4 acc = {0, 0, 0, 0};
    for (int it = 0; it < ITER; ++it) {
        const int base = ((it * 8 + w) * 1024 + lane * 16) % (64 * 1024);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            u4* p = reinterpret_cast<u4*>(L + ((base + r * 8192) % (64 * 1024)));
            if (MODE == 0) acc ^= *p;
            else *p = acc + (unsigned)r;
        }
        __asm__ volatile("" ::: "memory");
    }
    __syncthreads();
    out[blockIdx.x * 512 + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w ^ L[threadIdx.x];
}

// MODE 0: global_load_dwordx4 into VGPRs; 1: global_load_lds_dwordx4 into LDS.
template <int MODE>
__global__ __launch_bounds__(512) void l2_kernel(const unsigned char* __restrict__ g, unsigned* out) {
    __shared__ __attribute__((aligned(16))) unsigned char L[LDS_BYTES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    u4 acc = {0, 0, 0, 0};
    const int start = (blockIdx.x * 37) % CHUNKS;
    for (int p = 0; p < PASSES; ++p) {
        for (int c = w; c < CHUNKS; c += 64) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int ch = (start + c + u * 8) % CHUNKS;
                const unsigned char* src = g + (size_t)ch * 1024 + lane * 16;
                if (MODE == 0) acc ^= *reinterpret_cast<const u4*>(src);
                else __builtin_amdgcn_global_load_lds((gbl_ptr)src, (lds_ptr)(L + (w * 8 + u) * 1024), 16, 0, 0);
            }
            if (MODE == 1) __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        }
    }
    __syncthreads();
    out[blockIdx.x * 512 + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w ^ L[threadIdx.x];
}

}  // namespace

int main() {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return 1;
    unsigned* out = nullptr;
    unsigned char* g = nullptr;
    if (hipMalloc(&out, (size_t)cus * 512 * 4) != hipSuccess || hipMalloc(&g, BUF) != hipSuccess ||
        hipMemset(g, 1, BUF) != hipSuccess)
        return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    auto timed = [&](auto launch) -> float {
        launch();
        if (hipEventRecord(e0) != hipSuccess) return -1.0f;
        launch();
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.0f;
        float ms = 0;
        return hipEventElapsedTime(&ms, e0, e1) == hipSuccess ? ms : -1.0f;
    };
    const double lds_bytes = 8.0 * ITER * 16 * 1024;   // per workgroup
    const double l2_bytes = (double)PASSES * BUF;
    struct R {
        const char* name;
        float ms;
        double bytes;
    } rs[4] = {
        {"ds_read_b128", timed([&] { hipLaunchKernelGGL(lds_kernel<0>, dim3(cus), dim3(512), 0, 0, out); }), lds_bytes},
        {"ds_write_b128", timed([&] { hipLaunchKernelGGL(lds_kernel<1>, dim3(cus), dim3(512), 0, 0, out); }), lds_bytes},
        {"global_load_dwordx4 (L2)", timed([&] { hipLaunchKernelGGL(l2_kernel<0>, dim3(cus), dim3(512), 0, 0, g, out); }),
         l2_bytes},
        {"global_load_lds_dwordx4 (L2)",
         timed([&] { hipLaunchKernelGGL(l2_kernel<1>, dim3(cus), dim3(512), 0, 0, g, out); }), l2_bytes},
    };
    for (const R& r : rs) {
        if (r.ms <= 0) return 1;
        printf("%-30s %8.3f ms  %7.1f B/ns per CU  (%d CUs)\n", r.name, r.ms, r.bytes / (r.ms * 1e6), cus);
    }
    return (hipFree(out) == hipSuccess && hipFree(g) == hipSuccess) ? 0 : 1;
}
