// hbm_copy_probe.hip -- HBM reference kernels of known traffic:
//   copy   : 1 GiB read + 1 GiB written, 16 B per lane (calibrates rocprofv3's
//            FETCH_SIZE / WRITE_SIZE, KiB per dispatch, against true bytes)
//   read   : 4 GiB streamed, 16 B per lane, nothing written but one word per wave
//   gather : the access pattern of the HyperBall union at p = 7: a 4 B index read,
//            then one random 128 B record of a 4 GiB table (8 lanes x 16 B), for
//            2^27 indices -- the practical ceiling for the union kernels; also with
//            64 B and 96 B records packed back to back (what p = 6 needs, and what
//            p = 7 would need with 6-bit registers)
//   window : 16 B and 128 B records gathered from tables of 64 MiB .. 4 GiB, to show
//            where the working set stops fitting the memory-side cache
//   atomics: one 64-bit atomic add per block (or per wavefront) to the same address,
//            to size the cost of per-block counters
//   xcd    : (argument "xcd" runs only this) random 128 B records from a hot table of
//            8 parts; every block draws from the whole table, or block b only from part
//            b % 8, or only from part (b / 8) % 8 -- if workgroups go to the 8 XCDs
//            round-robin, the second keeps each XCD's L2 on one part
// Each is timed with HIP events, best of 5, over a few grid sizes.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

__global__ void hbm_copy(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void hbm_read(const uint4 *__restrict__ a, size_t n, uint32_t *out)
{
    uint32_t acc = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint4 v = a[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9e3779b9u) out[blockIdx.x] = acc;  // keeps the loads; never true for the fill pattern
}

__global__ void make_index(uint32_t *idx, size_t n, uint32_t records)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t z = (uint64_t)i + 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        idx[i] = (uint32_t)(z % records);
    }
}

// records of `units` x 16 B (units <= 8) at a stride of `units` x 16 B; a group of 8
// lanes serves one index, lanes >= units idle
__global__ void hbm_gather(const uint4 *__restrict__ table, const uint32_t *__restrict__ idx, size_t n,
                           uint32_t units, uint32_t *out)
{
    const uint32_t sub = threadIdx.x & 7u;
    uint32_t acc = 0;
    const size_t groups = (size_t)gridDim.x * blockDim.x / 8;
    for (size_t g = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / 8; g < n; g += groups) {
        const uint32_t r = idx[g];
        if (sub < units) {
            const uint4 v = table[(size_t)r * units + sub];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    if (acc == 0x9e3779b9u) out[blockIdx.x] = acc;
}

__global__ void xcd_gather(const uint4 *__restrict__ table, const uint32_t *__restrict__ idx, size_t n,
                           uint32_t per_part, uint32_t parts, int mode, uint32_t *out)
{
    const uint32_t sub = threadIdx.x & 7u;
    const uint32_t part = mode == 1 ? blockIdx.x % parts : (blockIdx.x / parts) % parts;
    uint32_t acc = 0;
    const size_t groups = (size_t)gridDim.x * blockDim.x / 8;
    for (size_t g = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / 8; g < n; g += groups) {
        const uint32_t r0 = idx[g];  // < per_part * parts
        const uint32_t r = mode == 0 ? r0 : part * per_part + r0 % per_part;
        const uint4 v = table[(size_t)r * 8 + sub];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9e3779b9u) out[blockIdx.x] = acc;
}

__global__ void atomic_per_block(unsigned long long *p, int per_wave)
{
    if (per_wave ? (threadIdx.x & 63) == 0 : threadIdx.x == 0) atomicAdd(p, 1ull);
}

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 1;                                                             \
        }                                                                         \
    } while (0)

template <class F>
static float best_ms(hipEvent_t e0, hipEvent_t e1, F launch)
{
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        (void)hipEventRecord(e0);
        launch();
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    return best;
}

int main(int argc, char **argv)
{
    const bool xcd_only = argc > 1 && argv[1][0] == 'x';
    const bool copy_only = argc > 1 && !xcd_only;  // the counter-calibration runs time the copy kernel alone
    const size_t copy_bytes = size_t(1) << 30, copy_n = copy_bytes / 16;
    const size_t table_bytes = size_t(4) << 30, table_n = table_bytes / 16;
    const uint32_t records = (uint32_t)(table_bytes / 128);
    const size_t n_idx = size_t(1) << 27;
    uint4 *a = nullptr, *b = nullptr, *table = nullptr;
    uint32_t *idx = nullptr, *sink = nullptr;
    CHECK(hipMalloc(&a, copy_bytes));
    CHECK(hipMalloc(&b, copy_bytes));
    CHECK(hipMalloc(&sink, 131072 * 4));
    CHECK(hipMemset(a, 1, copy_bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    if (xcd_only) {
        const size_t n = size_t(1) << 27;
        CHECK(hipMalloc(&table, size_t(1) << 30));
        CHECK(hipMalloc(&idx, n * 4));
        CHECK(hipMemset(table, 3, size_t(1) << 30));
        for (uint32_t mib : {16u, 32u, 64u, 128u}) {
            const uint32_t parts = 8, per_part = (mib << 20) / 128 / parts;  // 128 B records, < 1 GiB
            make_index<<<8192, 256>>>(idx, n, per_part * parts);
            CHECK(hipGetLastError());
            for (int mode : {0, 1, 2}) {
                const float ms = best_ms(e0, e1, [&] {
                    xcd_gather<<<16384, 256>>>(table, idx, n, per_part, parts, mode, sink);
                });
                CHECK(hipGetLastError());
                printf("xcd: hot table %u MiB, %s: %.3f ms, %.1f G records/s\n", mib,
                       mode == 0 ? "every block, whole table" : mode == 1 ? "block b, part b % 8" : "block b, part (b/8) % 8",
                       ms, n / (ms * 1e-3) / 1e9);
            }
        }
        CHECK(hipFree(table));
        CHECK(hipFree(idx));
        CHECK(hipFree(a));
        CHECK(hipFree(b));
        CHECK(hipFree(sink));
        return 0;
    }
    for (int grid : {8192, 32768, 131072}) {
        const float ms = best_ms(e0, e1, [&] { hbm_copy<<<grid, 256>>>(a, b, copy_n); });
        CHECK(hipGetLastError());
        printf("copy 1 GiB grid %d: %.3f ms, %.1f GB/s read+write\n", grid, ms, 2.0 * copy_bytes / (ms * 1e-3) / 1e9);
        if (copy_only) break;
    }
    if (!copy_only) {
        CHECK(hipMalloc(&table, table_bytes));
        CHECK(hipMalloc(&idx, n_idx * 4));
        CHECK(hipMemset(table, 3, table_bytes));
        make_index<<<8192, 256>>>(idx, n_idx, records);
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        for (int grid : {4096, 16384, 65536}) {
            const float ms = best_ms(e0, e1, [&] { hbm_read<<<grid, 256>>>(table, table_n, sink); });
            CHECK(hipGetLastError());
            printf("read 4 GiB grid %d: %.3f ms, %.1f GB/s\n", grid, ms, table_bytes / (ms * 1e-3) / 1e9);
        }
        for (uint32_t units : {8u, 6u, 4u}) {
            // every record index < records and units <= 8 keeps r * units + sub < table_n
            for (int grid : {4096, 16384, 65536}) {
                const float ms = best_ms(e0, e1, [&] { hbm_gather<<<grid, 256>>>(table, idx, n_idx, units, sink); });
                CHECK(hipGetLastError());
                const double bytes = (double)n_idx * (4 + 16 * units);
                printf("gather 2^27 x %u B (+4 B index) grid %d: %.3f ms, %.1f GB/s\n", 16 * units, grid, ms,
                       bytes / (ms * 1e-3) / 1e9);
            }
        }
        for (uint32_t units : {1u, 8u}) {
            for (size_t mib : {64, 128, 256, 512, 4096}) {
                const uint32_t recs = (uint32_t)((mib << 20) / (16 * units));  // records * units * 16 B <= table_bytes
                make_index<<<8192, 256>>>(idx, n_idx, recs);
                CHECK(hipGetLastError());
                const float ms = best_ms(e0, e1, [&] { hbm_gather<<<16384, 256>>>(table, idx, n_idx, units, sink); });
                CHECK(hipGetLastError());
                const double bytes = (double)n_idx * (4 + 16 * units);
                printf("window %zu MiB, %u B records: %.3f ms, %.1f GB/s (record bytes + index)\n", mib, 16 * units, ms,
                       bytes / (ms * 1e-3) / 1e9);
            }
        }
        unsigned long long *cnt = (unsigned long long *)sink;
        for (int per_wave : {0, 1}) {
            for (int grid : {1024, 8192, 32768}) {
                const float ms = best_ms(e0, e1, [&] { atomic_per_block<<<grid, 256>>>(cnt, per_wave); });
                CHECK(hipGetLastError());
                const int n_atomics = grid * (per_wave ? 4 : 1);
                printf("atomics: %d same-address u64 adds (%s), grid %d: %.3f ms, %.1f ns each\n", n_atomics,
                       per_wave ? "one per wavefront" : "one per block", grid, ms, ms * 1e6 / n_atomics);
            }
        }
        CHECK(hipFree(table));
        CHECK(hipFree(idx));
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(sink));
    return 0;
}
