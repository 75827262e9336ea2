// cupic_measure.hip -- plain streaming kernels of known traffic: the HBM
// bandwidth a copy, a read and a write reach on this device, and the
// reference point that scales profiler byte counters (FETCH_SIZE,
// WRITE_SIZE) into bytes.
#include <string>

#include "cupic.h"
#include "cupic_launch.h"

namespace cupic {
namespace {

typedef float v4 __attribute__((ext_vector_type(4)));

// one 16-byte element per thread: dst = src
__global__ __launch_bounds__(kBlock) void k_copy_probe(const v4 *__restrict__ src,
                                                       v4 *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// reads every element; one store per block keeps the loads live
__global__ __launch_bounds__(kBlock) void k_read_probe(const v4 *__restrict__ src,
                                                       float *__restrict__ out, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    float a = 0.0f;
    if (i < n) {
        const v4 v = src[i];
        a = (v.x + v.y) + (v.z + v.w);
    }
    a = wave_max_f(a);
    if ((threadIdx.x & (kWave - 1)) == 0 && a == -1.0f) out[blockIdx.x] = a;  // not taken: src holds ones
}

__global__ __launch_bounds__(kBlock) void k_write_probe(v4 *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = v4{1.0f, 1.0f, 1.0f, 1.0f};
}

}  // namespace
}  // namespace cupic

extern "C" int cupic_bandwidth_probe(int64_t bytes, int32_t kind, int32_t reps, double *ms)
{
    using namespace cupic;
    if (bytes < 16 || kind < 0 || kind > 2 || reps < 1 || !ms)
        return set_global_error(CUPIC_ERR_ARG, "bandwidth probe: bad argument");
    const long long n = bytes / 16;
    const long long nb = (n + kBlock - 1) / kBlock;
    if (nb > 0x7FFFFFFFLL) return set_global_error(CUPIC_ERR_ARG, "bandwidth probe: too large");
    v4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string msg;
    hipError_t e = hipMalloc(&a, (size_t)n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)n * 16);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto launch = [&]() {
        if (kind == 0)
            hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)nb), dim3(kBlock), 0, 0, a, b, n);
        else if (kind == 1)
            hipLaunchKernelGGL(k_read_probe, dim3((unsigned)nb), dim3(kBlock), 0, 0, a,
                               (float *)b, n);
        else
            hipLaunchKernelGGL(k_write_probe, dim3((unsigned)nb), dim3(kBlock), 0, 0, b, n);
    };
    if (e == hipSuccess) {
        // a source of ones (the read probe's store stays untaken)
        hipLaunchKernelGGL(k_write_probe, dim3((unsigned)nb), dim3(kBlock), 0, 0, a, n);
        launch();  // warm-up
        e = hipEventRecord(e0, 0);
        for (int r = 0; r < reps && e == hipSuccess; r++) launch();
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess) *ms = (double)t / reps;
    }
    if (e != hipSuccess) msg = std::string("bandwidth probe: ") + hipGetErrorString(e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (!msg.empty()) return set_global_error(CUPIC_ERR_HIP, msg.c_str());
    return CUPIC_OK;
}
