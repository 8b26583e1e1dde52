// LDS read-bandwidth microbenchmark: the peak the engine's LDS roofline is
// quoted against.  Every thread of 8 workgroups per CU reads 16 bytes per
// iteration from LDS at conflict-free consecutive addresses; prints the
// aggregate rate of three timed launches.  Built by build.py next to the
// engine library; run as `cpm-rust_amd/cpm/lds_bandwidth`.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                              \
    do {                                                                      \
        const hipError_t err_ = (x);                                          \
        if (err_ != hipSuccess) {                                             \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(err_));    \
            std::exit(1);                                                     \
        }                                                                     \
    } while (0)

__global__ void __launch_bounds__(256) k_lds_read(uint4* out, int iters) {
    __shared__ uint4 buf[2048];  // 32 KiB
    const int t = threadIdx.x;
    for (int i = t; i < 2048; i += 256) buf[i] = make_uint4(i, i + 1, i + 2, i + 3);
    __syncthreads();
    uint4 acc = make_uint4(0, 0, 0, 0);
    int idx = t;
    for (int it = 0; it < iters; ++it) {
        asm volatile("" ::: "memory");  // a fresh load every iteration
        const uint4 v = buf[idx];
        acc.x ^= v.x;
        acc.y += v.y;
        acc.z ^= v.z;
        acc.w += v.w;
        idx = (idx + 256) & 2047;
    }
    if (acc.x == 0x12345678u) out[blockIdx.x * 256 + t] = acc;  // keeps the loads
}

int main() {
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int blocks = prop.multiProcessorCount * 8, iters = 20000;
    uint4* out = nullptr;
    CHECK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint4)));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(256), 0, 0, out, 100);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    double best = 0.0;
    for (int r = 0; r < 3; ++r) {
        CHECK(hipEventRecord(a));
        hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(256), 0, 0, out, iters);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(b));
        CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, a, b));
        const double tbps = (double)blocks * 256 * 16 * iters / ms / 1e9;
        best = tbps > best ? tbps : best;
        std::printf("CUs %d, %d workgroups of 256: LDS read %.1f TB/s (%.3f ms)\n",
                    prop.multiProcessorCount, blocks, tbps, ms);
    }
    std::printf("lds_read_tbps: %.2f\n", best);
    CHECK(hipFree(out));
    return 0;
}
