// hb_kernels.h -- launchers of the HyperBall kernels (hb_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hb_device.h"
#include "../../include/sparkx_hb.h"

struct hb_mlp;

namespace hb {

// Degree-binned schedule of one run (device pointers).
struct HbSchedule {
    int nbins = 0;
    const uint32_t *bin_list[4] = {};  // null: the bin is the id range bin_base .. + bin_count
    uint32_t bin_base[4] = {};
    uint32_t bin_count[4] = {};
    int bin_team[4] = {};              // lanes per node in this bin
    const uint32_t *task_node = nullptr, *task_beg = nullptr, *task_end = nullptr;
    uint32_t ntasks = 0;               // hub neighbour chunks
    const uint32_t *hub_node = nullptr, *hub_task0 = nullptr, *hub_ntask = nullptr;
    uint32_t nhubs = 0;
    unsigned max_blocks = 8192;        // grid cap of the persistent loops
};

// Bin b is run by k_union_rows in its dense iterations (iteration 1 included) and by
// k_union in the others: the chain layout at p = 5 or 6, an id range from a multiple of
// 64, one sketch copy per node.
inline bool rows_bin(const HbDev &d, const HbSchedule &sc, int b)
{
    return d.rows && !sc.bin_list[b] && (sc.bin_base[b] & 63u) == 0 &&
           sc.bin_team[b] == (int)((1u << d.p) / 16);
}

// thread-local message returned by hb_last_error() (set by every C-ABI entry point)
void set_error(const char *msg);
// host-built HLL++ bias windows of precision p (4..16), built once per process
const HbBiasTable *bias_table(int p);

hipError_t launch_init(const HbDev &d, hipStream_t s);
hipError_t launch_fill(uint32_t *p, uint64_t count, uint32_t value, hipStream_t s);
// ev (nullable): 2 * HB_NSLOT events recorded around each kernel slot that launches
hipError_t launch_iteration(const HbDev &d, const HbSchedule &sc, bool nf, hipStream_t s,
                            hipEvent_t *ev = nullptr);
hipError_t launch_estimates(const HbDev &d, const uint8_t *regs, float *out, hipStream_t s);
// arc fingerprints into work[0..3] (work holds 4 + 4 * kFingerprintBlocks values):
// the CSR is symmetric iff work[0] == work[1] and work[2] == work[3]
constexpr uint32_t kFingerprintBlocks = 8192;
hipError_t arc_fingerprints(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint64_t arcs,
                            unsigned long long *work, hipStream_t s);
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
// Chain layout: nodes ordered by (nearest seed, level) of a multi-source breadth-first
// search from about one linked node in 64 (pseudo-random, fixed), then unreached linked
// nodes and isolated nodes by id; same outputs as relabel_by_degree.  *levels receives
// the number of levels searched (at most max_levels).  Synchronises s.
hipError_t relabel_by_chain(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t arcs, uint32_t *perm,
                            uint32_t *sigma, uint32_t *roff, uint32_t *rnbr, uint32_t *rorig, uint32_t max_levels,
                            uint32_t *levels, hipStream_t s);
// dst row u (caller's order) = src row sigma[u]; row_bytes a multiple of 4
hipError_t unpermute_rows(const void *src, void *dst, const uint32_t *sigma, uint32_t n, size_t row_bytes,
                          hipStream_t s);

// Device CSR construction (hb_ingest.hip).  off / nbr are allocated with hipMalloc
// (the caller takes them); err_code != 0 marks a rejected input at err_index (an MTX
// body line, 0-based, or a COO entry): 7 not a number, 8 zero index, 9 index >= n.
struct CsrBuild {
    uint32_t *off, *nbr;
    uint64_t arcs, emitted, entries;
    uint64_t err_index;
    int err_code;
};
namespace ingest {
struct MtxHeader {
    uint64_t n, nnz, body, body_line, err_line;  // body: offset of the line after the size line
    int symmetric;
};
// HB_MTX_* of the header / size line (0 if both parse)
int mtx_header(const char *text, uint64_t len, MtxHeader *h);
hipError_t mtx_body_to_csr(const uint8_t *text, uint64_t len, uint32_t n, bool symmetric, CsrBuild *out,
                           hipStream_t s);
hipError_t coo_to_csr(const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t n, uint32_t flags,
                      CsrBuild *out, hipStream_t s);
}  // namespace ingest

// Walk-count surrogate inputs (hb_walk.hip).  Lists longer than kWalkHub arcs are cut
// into kWalkHub-arc chunks: chunk c = arcs [beg[c], end[c]); hub h = node[h] owns
// chunks [first[h], first[h + 1]) (device arrays).
constexpr uint32_t kWalkHub = 4096;
struct WalkHubs {
    const uint32_t *beg = nullptr, *end = nullptr, *node = nullptr, *first = nullptr;
    uint32_t chunks = 0, hubs = 0;
};
// passes 2 or 3; s: 4 n u64 (S_1, S_2, S_3, M_1, each n long); deg: n u32; part:
// 2 * chunks u64; x: n * 6 f32 (the inputs listed in hb_walk_features), written for
// nodes < x_nodes
hipError_t walk_inputs(const uint32_t *off, const uint32_t *nbr, uint32_t n, const WalkHubs &hubs, int passes,
                       uint64_t *s, uint32_t *deg, uint64_t *part, float *x, uint32_t x_nodes, hipStream_t st);
// per-graph standardisation of inputs x, MLP inference and isolated-last ranking (the
// surrogate tail after its inputs); nodes >= infer_nodes have no neighbour (their
// output is the isolated mark without running the network) and their rows of x are
// only read when standardized_x asks for the standardised inputs back in x
hipError_t launch_surrogate_infer(struct ::hb_mlp *n, const uint32_t *off, uint32_t nodes, uint32_t infer_nodes,
                                  bool standardized_x, float *x_dev, float *out_dev, hipStream_t s);

// surrogate tail on a staged graph: features from snapshots, MLP inference (nodes
// >= infer_nodes have no neighbour)
hipError_t launch_surrogate_tail(struct ::hb_mlp *n, const uint32_t *off, const float *snaps, uint32_t nodes,
                                 uint32_t infer_nodes, bool standardized_x, float *x_dev, float *out_dev,
                                 hipStream_t s);
int mlp_device(const struct ::hb_mlp *n);
// exact BFS sums of sources [s_begin, s_end) over a staged CSR (hb_exact.hip), h_off
// the host copy of its offsets; symmetric: the CSR holds v -> u for every u -> v (levels
// may then pull)
int exact_bfs(const uint32_t *off, const uint32_t *h_off, const uint32_t *nbr, uint32_t n, bool symmetric,
              uint64_t s_begin,
              uint64_t s_end, uint64_t *reach, uint64_t *total, double *inv_sum, hipStream_t s);

}  // namespace hb
