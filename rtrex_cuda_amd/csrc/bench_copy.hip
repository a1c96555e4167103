// Streaming copy kernel of known size, used to calibrate the HBM byte
// counters (FETCH_SIZE / WRITE_SIZE) that price the decomposition kernels:
// one launch reads `bytes` and writes `bytes`, 16-byte vector accesses.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {
__global__ void __launch_bounds__(256) k_stream_copy(const uint4* __restrict__ a,
                                                     uint4* __restrict__ b, uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
        b[i] = a[i];
}
}  // namespace

// Copies `bytes` (a multiple of 16) device-to-device `iters` times with
// k_stream_copy on `device`; *ms receives the mean time of one launch.
// Returns 0 on success, 1 on a HIP error.
extern "C" int rtrex_bench_copy(int device, int64_t bytes, int iters, double* ms) {
    if (bytes <= 0 || bytes % 16 || iters <= 0) return 1;
    if (hipSetDevice(device) != hipSuccess) return 1;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 1;
    const uint64_t n = (uint64_t)bytes / 16;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) goto out;
    if (hipStreamCreate(&s) != hipSuccess) goto out;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto out;
    if (hipMemsetAsync(a, 1, bytes, s) != hipSuccess) goto out;
    {
        const unsigned grid = 256u * 64u;
        k_stream_copy<<<grid, 256, 0, s>>>((const uint4*)a, (uint4*)b, n);  // warm
        if (hipEventRecord(e0, s) != hipSuccess) goto out;
        for (int i = 0; i < iters; ++i)
            k_stream_copy<<<grid, 256, 0, s>>>((const uint4*)a, (uint4*)b, n);
        if (hipEventRecord(e1, s) != hipSuccess) goto out;
        if (hipEventSynchronize(e1) != hipSuccess) goto out;
        float t = 0;
        if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) goto out;
        *ms = (double)t / iters;
        rc = 0;
    }
out:
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
