// bench_lds_atomics.hip -- throughput of the LDS atomic adds the deposits can
// use on this part (ds_add_f32, ds_add_f64, ds_add_u32, ds_add_u64, no return
// value), behind the choice of FP64 accumulators in the FP64 form and 32-bit
// fixed point in the FP32 form (gev_deposit.hip).  256 threads per workgroup
// each add K times into an LDS array of S elements (every index masked to
// S - 1) in three address patterns: distinct per lane (0), pseudo-random (1),
// eight lanes on one address (2).  Prints one line per type and pattern.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                                                 \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);          \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

constexpr int S = 2048, K = 256, NB = 8192, REP = 5;

template <typename T> __global__ __launch_bounds__(256) void k_lds_add(T *out, int mode)
{
    __shared__ T acc[S];
    for (int i = threadIdx.x; i < S; i += 256) acc[i] = (T)0;
    __syncthreads();
    const uint32_t t = threadIdx.x;
    uint32_t h = t * 2654435761u + blockIdx.x;
    for (int k = 0; k < K; k++) {
        uint32_t a;
        if (mode == 0) {
            a = (t * 8 + k * 37) & (S - 1);
        } else if (mode == 1) {
            h = h * 1664525u + 1013904223u;
            a = (h >> 8) & (S - 1);
        } else {
            a = ((t >> 3) + k * 37) & (S - 1);
        }
        __hip_atomic_fetch_add(&acc[a], (T)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < S; i += 256) out[(size_t)blockIdx.x * S + i] = acc[i];
}

template <typename T> static int run(const char *name)
{
    T *out = nullptr;
    CHECK(hipMalloc(&out, sizeof(T) * (size_t)NB * S));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int mode = 0; mode < 3; mode++) {
        hipLaunchKernelGGL(k_lds_add<T>, dim3(NB), dim3(256), 0, 0, out, mode);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e0));
        for (int r = 0; r < REP; r++) hipLaunchKernelGGL(k_lds_add<T>, dim3(NB), dim3(256), 0, 0, out, mode);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double ops = (double)REP * NB * 256.0 * K;
        std::printf("%s pattern %d: %.3f ms per launch, %.1f G atomic adds/s\n", name, mode, ms / REP,
                    ops / (ms * 1e-3) / 1e9);
    }
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    CHECK(hipFree(out));
    return 0;
}

int main()
{
    if (run<float>("ds_add_f32")) return 1;
    if (run<double>("ds_add_f64")) return 1;
    if (run<uint32_t>("ds_add_u32")) return 1;
    if (run<unsigned long long>("ds_add_u64")) return 1;
    return 0;
}
