/*
 * hb_oracle.h -- CPU restatement of the reference HyperBall path.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests check the HIP path against this library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (sparkx_amd/libsparkx_hb.so) never links or calls it.
 *
 * Reference files restated here (paths relative to the reference repo root):
 *   src/centralities/hyper_ball.rs:702-736   init: empty sketch + insert(u as usize)
 *   src/centralities/hyper_ball.rs:754-874   the iteration (double buffer, f32 diff,
 *                                            f64 accumulators, termination, buffer parity)
 *   src/centralities/hyper_ball.rs:434-555   closeness / harmonic / Lin epilogues
 *   src/graph/cache/parser_mtx/producers.rs:228-234, impl_sort.rs:10-51
 *                                            MTX symmetric emission and per-node sort
 *   src/graph/cache/parser_mtx/{mod.rs,producers.rs}
 *                                            MatrixMarket text -> CSR (mtx_oracle.c)
 *   src/trails/bfs.rs:199-234, src/test_common.rs:195-249
 *                                            exact closeness (source not pre-marked)
 *   src/utils.rs:57-100                      mae / mape / ranks / spearman_rho
 *   gen_dbg_mtx.py:4-22                      de Bruijn graph family used by the fixtures
 *
 * The HLL arithmetic itself lives in the third-party crate hyperloglog-rs
 * (crates.io 0.1.56, patched to git rev 77e0350), which is not available
 * offline.  Its unknown details are exposed as hbo_variant knobs; the pinned
 * choice (the one that reproduces the reference's published accuracy numbers
 * bit-for-bit) is HBO_VARIANT_PINNED below.
 */
#ifndef HB_ORACLE_H
#define HB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- HLL variant knobs ------------------------------------------------- */
typedef struct {
    int hash;      /* 0: SipHash-1-3 key (0,0); 1: SipHash-2-4 key (0,0)                */
    int index;     /* 0: idx = top p bits, rank from (h<<p)|1<<(p-1);
                      1: idx = low p bits, rank = leading zeros of h>>p within 64-p bits +1;
                      2: idx = low p bits, rank = trailing zeros of h>>p + 1 (capped)
                      3: idx = top p bits, rank from (h<<p) without sentinel, capped at 63 */
    int lc_rule;   /* 0: zeros>0 && LC <= HLL++ threshold(p) -> LC (checked first);
                      1: classic: raw <= 2.5m && zeros>0 -> LC                       */
    int lc_calc;   /* 0: LC = (float)(m * ln(m/z)) in f64;  1: all-f32 arithmetic          */
    int sum_order; /* 0: sequential over registers, f32, padding excluded;
                      1: per-u32-word f32 partials (5 regs/word), padding registers
                         included (2^-0 each), then padding count subtracted;
                      2: per-word f32 partials, padding excluded;
                      3: sequential over all 5*W slots incl. padding, then subtract;
                      4: f64 sum, cast to f32                                           */
    int alpha;     /* 0: f32 alpha (0.673/0.697/0.709, else 0.7213/(1+1.079/m) in f32);
                      1: same constants evaluated in f64 then cast                       */
    int final_div; /* 0: (alpha*m*m)/S ;  1: alpha*((m*m)/S) ; 2: alpha*m*(m/S)           */
    int bias;      /* 0: none; 1: HLL++ bias correction (k=6 nearest points) when E <= 5m */
} hbo_variant;

/* The variant that reproduces the reference's published numbers; filled in by
 * the pinning search (tests/golden/pin_hll_variant.py). */
extern const hbo_variant HBO_VARIANT_PINNED;

/* ---- hashing ------------------------------------------------------------ */
uint64_t hbo_siphash(uint64_t k0, uint64_t k1, const uint8_t *msg, size_t len,
                     int c_rounds, int d_rounds);
/* hash of a Rust `usize` value under variant v (8 little-endian bytes). */
uint64_t hbo_hash_usize(const hbo_variant *v, uint64_t x);
/* register index and rank that insert(x) sets; returns index, writes rank. */
uint32_t hbo_index_rank(const hbo_variant *v, int p, uint64_t x, uint8_t *rank);

/* ---- estimator ---------------------------------------------------------- */
/* f32 cardinality estimate of m = 2^p u8 registers (values 0..63). */
float hbo_estimate_u8(const hbo_variant *v, int p, const uint8_t *regs);
/* same for the reference's packed layout: 5 six-bit registers per u32 word. */
float hbo_estimate_packed(const hbo_variant *v, int p, const uint32_t *words);
/* HLL++ bias of the raw estimate e at precision p (mean of the 6 nearest points). */
double hbo_knn_bias(int p, double e);
void hbo_knn_bias_many(int p, const float *e, uint64_t n, double *out);
/* pack / unpack between u8 registers and 6-bit packed words (ceil(m/5) words). */
void hbo_pack6(int p, const uint8_t *regs, uint32_t *words);
void hbo_unpack6(int p, const uint32_t *words, uint8_t *regs);

/* ---- HyperBall ---------------------------------------------------------- */
typedef void (*hbo_iter_cb)(void *user, int t, const uint8_t *regs_out,
                            const double *dist, const double *inv);

typedef struct {
    int iterations;          /* iterations executed, incl. the final zero-increase one */
    int final_parity_latest; /* 1 if the reference's counters file holds the last
                                iteration's output, 0 if it holds its input        */
    uint64_t increases[1024];/* per-iteration count of nodes whose estimate increased */
    double init_seconds;     /* wall time of the register initialisation */
    double loop_seconds;     /* wall time of the iterations               */
    uint64_t regime[3];      /* estimates after the union, over all (node, iteration):
                                [0] linear counting, [1] raw with the HLL++ bias
                                subtracted (raw <= 5m), [2] raw as is */
    double iter_seconds[1024]; /* wall time of each iteration (all threads) */
} hbo_stats;

/*
 * Runs HyperBall on a CSR (offsets u64[n+1], neighbours u32[arcs]).
 *   threads   >= 1 : reference partitioning (ceil(n/T) contiguous blocks);
 *   packed    != 0 : faithful 6-bit packed registers (the reference's memory layout),
 *                    else u8 registers; both give identical results.
 *   max_iters == 0 : run to the reference's convergence rule; > 0 : truncate.
 * Outputs (caller-allocated):
 *   regs_ref   n*2^p bytes: the buffer the reference leaves in its counters file
 *   regs_last  n*2^p bytes: output of the last executed iteration (may be NULL)
 *   dist, inv  f64[n] accumulators
 * Returns 0 on success.
 */
int hbo_hyperball(const uint64_t *offsets, const uint32_t *nbrs, uint64_t n, int p,
                  const hbo_variant *v, int threads, int packed, int max_iters,
                  uint8_t *regs_ref, uint8_t *regs_last, double *dist, double *inv,
                  hbo_stats *stats, hbo_iter_cb cb, void *user);

/* Single-thread timing of the first iteration's node work on a pseudo-random sample
 * of about one node in `stride` (registers initialised first, not timed);
 * seconds[reps] per repetition, *nodes / *arcs the sample's size. */
int hbo_node_sample(const uint64_t *offsets, const uint32_t *nbrs, uint64_t n, int p, const hbo_variant *v,
                    int packed, uint64_t stride, int reps, double *seconds, uint64_t *nodes, uint64_t *arcs);

/* estimates of every node's sketch (u8 registers) */
void hbo_estimates(const hbo_variant *v, int p, uint64_t n, const uint8_t *regs, float *out);

/* epilogues: kind 0 closeness, 1 harmonic, 2 lin; normalize as the reference flag */
void hbo_centrality(int kind, int normalize, uint64_t n, const float *est,
                    const double *dist, const double *inv, double *out);

/* ---- exact BFS (reference quirk: source not pre-marked visited) ---- */
/* per source in [src_begin, src_end): reachable count, the f64 distance sum, and
 * (inv_sum may be NULL) the node-order sum of 1/d */
void hbo_exact_bfs(const uint64_t *offsets, const uint32_t *nbrs, uint64_t n, uint64_t src_begin,
                   uint64_t src_end, uint64_t *reach, double *total, double *inv_sum);
void hbo_exact_closeness(const uint64_t *offsets, const uint32_t *nbrs, uint64_t n,
                         uint64_t src_begin, uint64_t src_end, double *out);

/* ---- metrics (reference semantics) ------------------------------------- */
double hbo_mae(const double *a, const double *b, uint64_t n);
double hbo_mape(const double *a, const double *b, uint64_t n);
double hbo_spearman(const double *a, const double *b, uint64_t n);
void hbo_ranks(const double *v, uint64_t n, uint64_t *rank_out);

/* ---- fixture graphs ------------------------------------------------------ */
/* gen_dbg_mtx.py(sigma, k) parsed with the reference's MTX rules: symmetric
 * emission (self-loop once, duplicates kept), per-node ascending sort.
 * Returns arcs; offsets must hold n+1, nbrs may be NULL to query the size. */
uint64_t hbo_dbg_mtx_csr(int sigma, int k, uint64_t *offsets, uint32_t *nbrs);

/* ---- MatrixMarket ingestion (mtx_oracle.c) ------------------------------ */
typedef struct {
    uint64_t n, nnz;      /* size line: rows (= columns), declared entries              */
    uint64_t body;        /* byte offset of the first line after the size line          */
    uint64_t body_line;   /* its 1-based line number                                    */
    uint64_t entries;     /* entry lines used                                           */
    uint64_t arcs;        /* arcs emitted                                               */
    uint64_t err_line;    /* 1-based line of the first error                            */
    int symmetric;        /* "symmetric" or "skew-symmetric": entries emit both ways    */
} hbo_mtx_info;

enum {
    HBO_MTX_BAD_HEADER = -1,       /* first line does not start with %%MatrixMarket      */
    HBO_MTX_MALFORMED_HEADER = -2, /* fewer than 5 header tokens                         */
    HBO_MTX_UNSUPPORTED = -3,      /* not matrix coordinate pattern|integer sym|skew|gen */
    HBO_MTX_NO_SIZE_LINE = -4,     /* end of text before the size line                   */
    HBO_MTX_BAD_SIZE_LINE = -5,    /* size line missing a number or not a number         */
    HBO_MTX_NOT_SQUARE = -6,       /* nrows != ncols                                     */
    HBO_MTX_BAD_ENTRY = -7,        /* an index token is not a usize                      */
    HBO_MTX_ZERO_INDEX = -8,       /* 0 index (MatrixMarket is 1-based)                  */
    HBO_MTX_OUT_OF_BOUNDS = -9,    /* index > n                                          */
    HBO_MTX_NOMEM = -10,
};

/* Header only (fills n, nnz, body, body_line, symmetric). */
int hbo_mtx_header(const char *text, uint64_t len, hbo_mtx_info *info);
/* Whole text -> CSR: offsets[n+1] (may be NULL to get sizes), nbrs[arcs] (may be NULL).
 * Returns 0 or an HBO_MTX_* code with info->err_line set. */
int hbo_parse_mtx(const char *text, uint64_t len, hbo_mtx_info *info, uint64_t *offsets, uint32_t *nbrs);

#ifdef __cplusplus
}
#endif
#endif
