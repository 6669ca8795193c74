// hb_kernels.h -- launchers of the HyperBall kernels (hb_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hb_device.h"

struct hb_mlp;

namespace hb {

// Degree-binned schedule of one run (device pointers).
struct HbSchedule {
    int nbins = 0;
    const uint32_t *bin_list[4] = {};
    uint32_t bin_count[4] = {};
    int bin_team[4] = {};              // lanes per node in this bin
    const uint32_t *task_node = nullptr, *task_beg = nullptr, *task_end = nullptr;
    uint32_t ntasks = 0;               // hub neighbour chunks
    const uint32_t *hub_node = nullptr, *hub_task0 = nullptr, *hub_ntask = nullptr;
    uint32_t nhubs = 0;
    unsigned max_blocks = 8192;        // grid cap of the persistent loops
};

// thread-local message returned by hb_last_error() (set by every C-ABI entry point)
void set_error(const char *msg);

hipError_t launch_init(const HbDev &d, hipStream_t s);
hipError_t launch_fill(uint32_t *p, uint64_t count, uint32_t value, hipStream_t s);
// ev (nullable): 2 * HB_NSLOT events recorded around each kernel slot that launches
hipError_t launch_iteration(const HbDev &d, const HbSchedule &sc, bool nf, hipStream_t s,
                            hipEvent_t *ev = nullptr);
hipError_t launch_estimates(const HbDev &d, const uint8_t *regs, float *out, hipStream_t s);
// arc fingerprints into work[0..3] (work holds 4 + 4 * kFingerprintBlocks values):
// the CSR is symmetric iff work[0] == work[1] and work[2] == work[3]
constexpr uint32_t kFingerprintBlocks = 8192;
hipError_t arc_fingerprints(const uint32_t *off, const uint32_t *nbr, uint32_t n, unsigned long long *work,
                            hipStream_t s);
// out[k] = sum_r part[r * cols + k] in a fixed order (T = unsigned long long or double)
template <typename T>
hipError_t sum_columns(const T *part, uint32_t rows, uint32_t cols, T *out, hipStream_t s);
hipError_t launch_centrality(const HbDev &d, const float *est, int kind, int normalize, double *out,
                             hipStream_t s);

// Degree-ordered layout (hb_layout.hip): perm[i] = caller's id of node i, sigma its
// inverse; roff / rnbr the renumbered CSR (nodes by descending degree, ties by id;
// each list ascending), rorig[e] the caller's id of rnbr[e].  Synchronises s.
hipError_t relabel_by_degree(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t arcs, uint32_t *perm,
                             uint32_t *sigma, uint32_t *roff, uint32_t *rnbr, uint32_t *rorig, hipStream_t s);
// dst row u (caller's order) = src row sigma[u]; row_bytes a multiple of 4
hipError_t unpermute_rows(const void *src, void *dst, const uint32_t *sigma, uint32_t n, size_t row_bytes,
                          hipStream_t s);

// surrogate tail on a staged graph: features from snapshots, MLP inference
hipError_t launch_surrogate_tail(struct ::hb_mlp *n, const uint32_t *off, const float *snaps, uint32_t nodes,
                                 float *x_dev, float *out_dev, hipStream_t s);
int mlp_device(const struct ::hb_mlp *n);
// exact BFS sums of sources [s_begin, s_end) over a staged CSR (hb_exact.hip)
int exact_bfs(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint64_t s_begin, uint64_t s_end,
              uint64_t *reach, uint64_t *total, double *inv_sum, hipStream_t s);

}  // namespace hb
