// hbm_copy_probe.hip -- plain 1 GiB device-to-device copy kernel of known size.
// Used to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters (KiB per
// dispatch) against true bytes moved, and to measure streaming HBM bandwidth.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void hbm_copy(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main()
{
    const size_t bytes = size_t(1) << 30, n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int grid : {8192, 32768, 131072}) {
        CHECK(hipEventRecord(e0));
        hbm_copy<<<grid, 256>>>(a, b, n);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        printf("copy 1 GiB grid %d: %.3f ms, %.1f GB/s read+write\n", grid, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
