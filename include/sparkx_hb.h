/*
 * sparkx_hb.h -- C ABI of the MI355X HyperBall engine (libsparkx_hb.so).
 *
 * Drop-in boundary for the reference's HyperBall path: HyperBallInner<'a,N,E,Ix,P,6>
 * on &GraphMemoryMap<N,E,Ix> (src/centralities/hyper_ball.rs:81-108).  Plain
 * pointers and sizes only; every call returns an hb_status (0 = ok, < 0 = error)
 * and leaves a thread-local message for hb_last_error().  Host buffers are
 * caller-owned; a context owns its device memory and one HIP stream and is not
 * thread-safe (use one context per host thread).
 */
#ifndef SPARKX_HB_H
#define SPARKX_HB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HB_OK = 0,
    HB_EINVAL = -1,   /* bad argument (null pointer, p out of 4..16, malformed CSR) */
    HB_ENOMEM = -2,   /* device or pinned host allocation failed                */
    HB_EHIP = -3,     /* a HIP runtime call failed                              */
    HB_ESTATE = -4,   /* call out of order (e.g. hb_run before hb_stage_csr)   */
    HB_ERANGE = -5,   /* node id / size beyond what the engine supports         */
} hb_status;

/* hb_run flags */
#define HB_FLAG_DENSE 0x1u    /* recompute every node every iteration (the reference's
                                 schedule); default skips nodes none of whose
                                 neighbours changed, with identical results      */
#define HB_FLAG_GRAPH 0x2u    /* capture one iteration in a hipGraph and replay it   */
#define HB_FLAG_NF    0x4u    /* record the neighbourhood function N(t) each iteration */
#define HB_FLAG_TIMING 0x8u   /* per-iteration device timing with events (read at the
                                 host's usual checks: no extra sync)             */
#define HB_FLAG_KTIMING 0x10u /* per-kernel device timing with events (implies TIMING;
                                 not combinable with HB_FLAG_GRAPH)              */

/* hb_read_registers / hb_read_estimates selector */
#define HB_REGS_REFERENCE 0   /* the buffer the reference leaves in its counters file:
                                 the last iteration's output if the iteration count is
                                 even, its input if odd (hyper_ball.rs:840,865)      */
#define HB_REGS_LATEST 1      /* output of the last executed iteration              */

/* hb_set_layout: node numbering on the device */
#define HB_LAYOUT_INPUT 0     /* the caller's numbering */
#define HB_LAYOUT_DEGREE 1    /* nodes by descending degree (ties by id), each neighbour
                                 list ascending; built on the device when the graph is
                                 staged.  Results are identical and every read returns
                                 the caller's order. */
#define HB_LAYOUT_CHAIN 2     /* nodes by (nearest seed, level) of a multi-source
                                 breadth-first search from about one node in 64: along
                                 long low-degree paths (de Bruijn graphs of genomes)
                                 neighbours get nearby numbers and share cache lines */
#define HB_LAYOUT_AUTO 3      /* default: CHAIN when every degree is <= 64 and there are
                                 at most 4 arcs per node on average, DEGREE otherwise */

/* hb_centrality kinds */
#define HB_CLOSENESS 0
#define HB_HARMONIC 1
#define HB_LIN 2

#define HB_STATS_MAX_ITERS 4096
/* kernel slots of one iteration: union over degree bins 0..2, hub chunks, hub merge */
#define HB_NKERNELS 5

typedef struct {
    int32_t iterations;          /* executed, including the final zero-increase one   */
    int32_t converged;           /* 1: stopped by the reference's rule; 0: truncated  */
    int32_t final_parity_latest; /* 1 if HB_REGS_REFERENCE == HB_REGS_LATEST          */
    int32_t p;
    double ms_total;             /* device time of the iteration loop (events)         */
    double ms_init;              /* device time of the register initialisation        */
    uint64_t arcs_gathered;      /* sum over iterations of neighbour sketches gathered */
    uint64_t nodes_recomputed;   /* sum over iterations of nodes recomputed            */
    uint64_t increases[HB_STATS_MAX_ITERS];    /* nodes whose estimate increased, per t */
    uint64_t iter_arcs[HB_STATS_MAX_ITERS];
    uint64_t iter_nodes[HB_STATS_MAX_ITERS];
    float iter_ms[HB_STATS_MAX_ITERS];         /* filled with HB_FLAG_TIMING          */
    double nf[HB_STATS_MAX_ITERS];             /* N(t) = sum_u est_t(u), HB_FLAG_NF    */
    /* per kernel slot: launches, device ms (HB_FLAG_KTIMING), work done, schedule */
    uint64_t kernel_launches[HB_NKERNELS];
    double kernel_ms[HB_NKERNELS];
    uint64_t kernel_arcs[HB_NKERNELS];    /* neighbour (or partial) sketches gathered */
    uint64_t kernel_nodes[HB_NKERNELS];   /* nodes recomputed                         */
    uint32_t kernel_items[HB_NKERNELS];   /* nodes (bins, merge) or tasks (chunks) scheduled */
    uint32_t kernel_team[HB_NKERNELS];    /* lanes per node / task                    */
    /* the part of kernel_arcs / kernel_nodes done in iteration 1 with p <= 10, where
       every sketch is still a single insertion: the neighbour's (and the node's own)
       2^p-byte register block is not read, its id is hashed instead */
    uint64_t kernel_first_arcs[HB_NKERNELS];
    uint64_t kernel_first_nodes[HB_NKERNELS];
    /* iterations whose recomputed nodes were listed from the changed nodes' neighbour
       lists (symmetric graphs, few changes) instead of by scanning every node */
    int32_t push_iterations;
    int32_t symmetric;           /* the staged CSR is symmetric */
    /* neighbour register blocks actually read, per kernel slot: kernel_arcs counts the
       neighbour ids scanned; iteration 1 hashes ids instead of reading blocks, and
       outside dense iterations only neighbours that changed are read */
    uint64_t kernel_gathered[HB_NKERNELS];
    int32_t layout;              /* HB_LAYOUT_* the run used (INPUT, DEGREE or CHAIN) */
    float ms_layout;             /* device time it took to build the degree layout of the
                                    staged graph (at hb_stage_csr / hb_set_layout; not
                                    part of ms_init or ms_total); 0 for HB_LAYOUT_INPUT */
} hb_stats;

typedef struct hb_ctx hb_ctx;

/* Creates a context on HIP device `device_ordinal`.
 * Replaces: HyperBallInner::new_no_compute_impl's resource set-up
 * (src/centralities/hyper_ball.rs:684-700, :738-752). */
int hb_create(hb_ctx **out, int device_ordinal);

/* Releases the context and all its device memory.
 * Replaces: dropping HyperBallInner (its mmaps). */
void hb_destroy(hb_ctx *ctx);

/* Message of the last failing call on this thread ("" if none). */
const char *hb_last_error(void);

/* SHA-256 (hex) of the compile flags and the source files this library was built from,
 * as sparkx_amd/build.py computes it; "unknown" for a build made without it.  No
 * counterpart in the reference: it lets a caller or a test show which sources a loaded
 * binary comes from. */
const char *hb_build_digest(void);

/* Stages a CSR graph into HBM once, through pinned host buffers.  offsets has n+1
 * entries and nbrs has arcs entries, each 8- or 4-byte unsigned (the reference's
 * mmap'd index and edge files are 8-byte).  Neighbour ids must be < n and offsets
 * non-decreasing with offsets[0] == 0 and offsets[n] == arcs; n < 2^32, arcs < 2^32.
 * Replaces: the graph borrow g: &GraphMemoryMap and its CSR walk
 * GraphMemoryMap::{size,width,index_node,neighbours} (src/graph/mod.rs:569-595,648-658). */
int hb_stage_csr(hb_ctx *ctx, const void *offsets, int off_bytes, uint64_t n,
                 const void *nbrs, int nbr_bytes, uint64_t arcs);

/* Runs HyperBall with precision p (2^p registers per node, B = 6) on the staged
 * graph.  max_iters == 0 follows the reference: iterate until no node's estimate
 * increased; max_iters > 0 truncates (an extension: the reference has no cap).
 * Replaces: HyperBallInner::new / new_with_conf and compute_with_proc_mem_impl
 * (src/centralities/hyper_ball.rs:124-157, :702-736, :754-874). `stats` may be NULL. */
int hb_run(hb_ctx *ctx, int p, int max_iters, uint32_t flags, hb_stats *stats);

/* Rebuilds a completed run of precision p on the staged graph from saved results:
 * regs (n * 2^p u8 registers, node-major, the buffer HB_REGS_REFERENCE reads) and the
 * f64 accumulators dist / inv, all in the caller's node order.  Every reader and
 * epilogue then works as after hb_run; the restored run reports 0 iterations.
 * Replaces: HyperBallInner::get_or_compute reopening the counters and accumulator
 * files (src/centralities/hyper_ball.rs:165-201; its restored struct has iters = 0). */
int hb_restore(hb_ctx *ctx, int p, const uint8_t *regs, const double *dist, const double *inv);

/* Chooses the device numbering of the nodes (HB_LAYOUT_*) for the following runs
 * (builds it now if a graph is staged); discards the results of an earlier run if it
 * changes.  No reference counterpart: the reference walks the caller's CSR as given. */
int hb_set_layout(hb_ctx *ctx, int layout);

/* Runs hb_run(ctxs[i], p[i], max_iters, flags, stats + i) for i < count (stats may be
 * NULL), each context on its own stream, from a pool of `threads` host threads (0: up
 * to 8), so the GPU interleaves many independent graphs.  Every context runs to its
 * own end; the first failing context (by index) is reported.  Contexts must be
 * distinct.  The batched form of HyperBallInner::new over many graphs (no reference
 * counterpart beyond calling it once per graph). */
int hb_run_many(hb_ctx **ctxs, const int *p, int count, int max_iters, uint32_t flags,
                hb_stats *stats, int threads);

/* Copies n * 2^p u8 registers (node-major) of the selected buffer to dst.
 * Replaces: the counters file HyperBallInner::counters (hyper_ball.rs:92). */
int hb_read_registers(hb_ctx *ctx, int which, uint8_t *dst);

/* f32 estimate of every node's sketch in the selected buffer.
 * Replaces: HyperLogLog::estimate_cardinality() on counters (hyper_ball.rs:452,506,544). */
int hb_read_estimates(hb_ctx *ctx, int which, float *dst);

/* Copies the f64 distance and inverse-distance accumulators (either may be NULL).
 * Replaces: geodesic_distances / harmonic_geodesic_distances (hyper_ball.rs:273-280). */
int hb_read_accumulators(hb_ctx *ctx, double *dist, double *inv);

/* Closeness / harmonic / Lin centrality from the last run, computed on the device.
 * Replaces: compute_closeness_centrality / compute_harmonic_centrality /
 * compute_lins_centrality (hyper_ball.rs:434-555), with the reference's formulas. */
int hb_centrality(hb_ctx *ctx, int kind, int normalize, double *dst);

/* N(t) for t = 1..min(iterations, cap) of the last run made with HB_FLAG_NF;
 * returns the count written (>= 0) or an hb_status.  No reference counterpart. */
int hb_neighbourhood_function(hb_ctx *ctx, double *dst, int cap);

/* Scheduling knobs (results never depend on them).  Any 0 keeps the current value.
 *   small_max_deg : nodes up to this degree use one 2^p/16-lane group per node
 *   wave_max_deg  : nodes up to this degree use one wavefront per node
 *   chunk         : neighbours per wavefront task for nodes above wave_max_deg
 * No reference counterpart (the reference uses static contiguous thread blocks). */
int hb_set_schedule(hb_ctx *ctx, uint32_t small_max_deg, uint32_t wave_max_deg,
                    uint32_t chunk);

/* ---- batches: many independent graphs advanced together ------------------
 * No reference counterpart (the reference runs one graph per HyperBallInner).
 * Graphs of mixed size and precision (p = 4..8) are advanced together: one hipGraph
 * replay per iteration of persistent kernels over the (graph, degree-bin) tiles of the
 * graphs still running; each graph keeps its own iteration count, termination rule and
 * buffer parity, and its results are bit-identical to hb_run on that graph alone.
 * SPARKX_HB_TRACE=1 in the environment prints one progress line per host check. */
typedef struct {
    const void *offsets;
    int off_bytes;      /* 4 or 8 */
    uint64_t n;         /* 1 <= n < 2^31 */
    const void *nbrs;
    int nbr_bytes;      /* 4 or 8 */
    uint64_t arcs;
    int p;              /* 4..8 */
    int max_iters;      /* 0: the reference's convergence rule */
} hb_graph_desc;

typedef struct {
    int32_t launches;          /* iteration launches until every graph stopped */
    int32_t iterations_max;    /* longest graph run                            */
    uint64_t iterations_total; /* sum of every graph's iterations              */
    uint64_t arcs_gathered;
    uint32_t tiles;            /* work-queue tiles per iteration               */
    double ms_total;           /* device time of the iterations                */
    double ms_init;
    uint64_t nodes_recomputed;
    uint64_t alg_bytes;        /* algorithmic bytes of all iterations: per recomputed
                                  node 8 + 2*2^p + 40, per gathered arc 4 + 2^p     */
} hb_batch_stats;

typedef struct hb_batch hb_batch;

/* Validates and stages `count` graphs into HBM. */
int hb_batch_create(hb_batch **out, int device_ordinal, const hb_graph_desc *graphs, int count);
void hb_batch_destroy(hb_batch *b);
/* Tiling knobs (results never depend on them); 0 keeps the current value. */
int hb_batch_set_schedule(hb_batch *b, uint32_t small_max_deg, uint32_t mid_max_deg, uint32_t tile_arcs);
/* Runs every graph to its own stopping point; flags: HB_FLAG_DENSE, HB_FLAG_GRAPH. */
int hb_batch_run(hb_batch *b, uint32_t flags, hb_batch_stats *stats);
/* Iterations, convergence and per-iteration increase counts of one graph;
 * returns the number of increase counts written (<= cap). */
int hb_batch_graph_stats(hb_batch *b, int graph, int32_t *iterations, int32_t *converged,
                         uint64_t *increases, int cap);
int hb_batch_read_registers(hb_batch *b, int graph, int which, uint8_t *dst);
int hb_batch_read_estimates(hb_batch *b, int graph, int which, float *dst);
int hb_batch_read_accumulators(hb_batch *b, int graph, double *dist, double *inv);

/* ---- autotuning -----------------------------------------------------------
 * Times a grid of schedules on the first `trial_iters` iterations of this graph
 * (or batch) and keeps the fastest; results never depend on the choice.  Writes
 * the chosen knobs and their trial time.  No reference counterpart. */
int hb_autotune(hb_ctx *ctx, int p, int trial_iters, uint32_t *small_max_deg,
                uint32_t *wave_max_deg, uint32_t *chunk, double *trial_ms);
int hb_batch_autotune(hb_batch *b, uint32_t *small_max_deg, uint32_t *mid_max_deg,
                      uint32_t *tile_arcs, double *trial_ms);

/* ---- estimate snapshots ------------------------------------------------------
 * Record every node's estimate after iterations 1..k of the following runs
 * (k <= 16; 0 turns it off).  hb_read_snapshots copies k*n floats, iteration-major,
 * with k the count in force when the completed run started (hb_result_info reports
 * it; a later hb_set_snapshots changes the next run only); rows past the last
 * executed iteration repeat its estimates.  No reference counterpart (they are the
 * learned surrogate's inputs N_1..N_k). */
int hb_set_snapshots(hb_ctx *ctx, int k);
int hb_read_snapshots(hb_ctx *ctx, float *dst);

/* Describes the results the context holds (each pointer may be NULL): the precision
 * p of the completed run (hb_read_registers copies n * 2^p bytes), its iteration
 * count and its snapshot rows.  HB_ESTATE without a completed run.  hb_surrogate
 * replaces the run with its own 4 iterations at p_feat.  No reference counterpart
 * (the reference's P is a type parameter, hyper_ball.rs:81). */
int hb_result_info(const hb_ctx *ctx, int *p, int *iterations, int *snapshots);

/* ---- learned surrogate ---------------------------------------------------------
 * An MLP 6 -> 64 -> 64 -> 1 with GELU (BF16 MFMA matmuls, f32 accumulation, f32
 * master weights and Adam).  Parameters are 4673 floats: W1[6][64], b1[64],
 * W2[64][64], b2[64], W3[64], b3; normalisation is mean[6], std[6] of the inputs.
 * In the surrogate the inputs are a node's degree, log(1+degree) and log N_1..N_4
 * (N_t its estimate after t HyperBall iterations), each standardised by the graph's
 * mean and standard deviation over nodes with a neighbour; the target is log(1+H),
 * H the harmonic accumulator of the full run, standardised the same way.
 * No reference counterpart. */
typedef struct hb_mlp hb_mlp;
int hb_mlp_create(hb_mlp **out, int device_ordinal, uint64_t seed);
void hb_mlp_destroy(hb_mlp *m);
int hb_mlp_set_params(hb_mlp *m, const float *params, const float *norm);
int hb_mlp_get_params(hb_mlp *m, float *params, float *norm);
/* mean-squared-error gradient over `count` samples (x: count*6, y: count) */
int hb_mlp_gradient(hb_mlp *m, const float *x, const float *y, uint64_t count, float *grad, float *loss);
/* Adam training (minibatch, shuffled per epoch); sets the input normalisation from x */
int hb_mlp_train(hb_mlp *m, const float *x, const float *y, uint64_t count, int epochs, int batch,
                 float lr, uint64_t seed, float *epoch_loss);
int hb_mlp_predict(hb_mlp *m, const float *x, uint64_t count, float *y);
/* Anytime ranking on the staged graph: HyperBall at precision p_feat for 4
 * iterations, device-side standardised inputs, MLP inference; pred (n floats)
 * receives the predicted standardised log harmonic centrality (nodes without
 * neighbours: -3e38, last), features (n*6, optional) the standardised inputs;
 * ms = device time of register init, the 4 iterations, inputs and inference
 * (host set-up excluded).  Replaces this context's previous run. */
int hb_surrogate(hb_ctx *ctx, hb_mlp *m, int p_feat, float *pred, float *features, double *ms);

/* Walk-count inputs (alternative input sets for the same MLP, no HyperBall
 * iteration): S_0 = degree, S_k(u) the sum of S_{k-1}(v) over u's arcs u -> v (walks of
 * k+1 arcs from u; u64, saturating at 2^64-1, so exact and order-free), M_1(u) the
 * largest degree among u's heads (0 without arcs), L(x) = f64 log1p(x).  Per node, 6
 * f32 in the caller's order, each rounded once from f64:
 *   passes = 2: L(deg), L(S_1), L(S_2), L(M_1), L(S_2) - L(S_1), L(S_1) - L(deg)
 *   passes = 3: L(deg), L(S_1), L(S_2), L(S_3), L(M_1), L(S_3) - L(S_2)
 * (one gather pass per S_k).  ms = device time.  Leaves the context's HyperBall run in
 * place.  No reference counterpart. */
int hb_walk_features(hb_ctx *ctx, int passes, float *features, double *ms);

/* Anytime ranking from walk-count inputs: hb_walk_features, standardised per graph as
 * in hb_surrogate, then the MLP (which must have been trained on the same input set);
 * pred and features (standardised, may be NULL) as in hb_surrogate; ms = device time
 * of the passes, the inputs and inference.  Leaves the context's HyperBall run in
 * place.  No reference counterpart. */
int hb_surrogate_walks(hb_ctx *ctx, hb_mlp *m, int passes, float *pred, float *features, double *ms);

/* ---- exact BFS centralities ---------------------------------------------------
 * One breadth-first search per source over the staged graph, 64 sources per pass
 * (one bit lane each), following arcs forwards as the reference's BFSDists does
 * (trails/bfs.rs:199-234), including its quirk: the source is not marked visited, so
 * it is rediscovered and counted.  Sources are [src_begin, src_end) (a sample, or
 * 0..n for every node).
 *
 * hb_exact_bfs writes, per source (each pointer may be NULL): the reachable count
 * (BFSDists::recheable), the distance total (total_distances, as an integer; the
 * reference's f64 sum of integers equals it below 2^53) and the node-order f64 sum
 * of 1/d over v = 0..n-1 (exact.rs:386-396, bit-identical to the reference's
 * sequential sum; needs eccentricity <= 65535, else HB_ERANGE).
 * Replaces: BFSDists::new + recheable()/total_distances()/get_distances(). */
int hb_exact_bfs(hb_ctx *ctx, uint64_t src_begin, uint64_t src_end, uint64_t *reach,
                 uint64_t *total, double *inv_sum);

/* Exact centrality of sources [src_begin, src_end) into dst (src_end - src_begin
 * doubles); kind HB_CLOSENESS / HB_HARMONIC / HB_LIN; normalization -1 = None,
 * 0 = Some(false), 1 = Some(true) (ignored for Lin), as the reference's Option<bool>.
 * Replaces: ExactClosenessCentrality::new (exact.rs:21-30,138-226),
 * ExactHarmonicCentrality::new (:238-249,357-466), ExactLinCentrality::new
 * (:476-486,560-644), restricted to a source range. */
int hb_exact_centrality(hb_ctx *ctx, int kind, int normalization, uint64_t src_begin,
                        uint64_t src_end, double *dst);

/* ---- evaluation metrics, computed on `device` -----------------------------------
 * v, a, b and rank may be host or device pointers (unified addressing); n doubles
 * each (n uint32 for rank).
 * hb_eval_ranks: rank[i] = position of v[i] in a stable descending order (ties keep
 * index order; -0.0 equals +0.0), bit-identical to the reference's ranks().
 * Replaces: utils::ranks (src/utils.rs:76-85).
 *
 * hb_eval_metrics: out[0] = MAE(a, b), out[1] = MAPE(a, b) in percent (terms with
 * b[i] == 0 count 0), out[2] = Spearman rho(a, b) over those ranks, out[3] = the
 * top-k overlap |top_k(a) & top_k(b)| / k with k = max(1, ceil(top_frac * n)).  The
 * squared rank differences are summed exactly (128-bit integers, rounded once), the
 * MAE / MAPE sums in a fixed two-stage order; both stay within a few ulps of the
 * reference's sequential f64 sums.
 * Replaces: utils::mae / mape / spearman_rho (src/utils.rs:57-74,87-100).
 *
 * 1 <= n < 2^32; a NaN anywhere gives HB_EINVAL (the reference's order for it is
 * unspecified). */
/* hb_bias_windows: the HLL++ bias-correction windows of precision p (4..16) as the
 * kernels read them, built on the host; no device work.  No counterpart in the
 * reference: it exposes the tabulated form of hyperloglog-rs's k = 6 nearest-neighbour
 * bias rule (estimate_cardinality, called at hyper_ball.rs:806,818) so that it can be
 * checked against that rule on the CPU.  mean[195]: mean bias of window s (points
 * s..s+5); breakpoints[195]: breakpoints[s] (s >= 1) is the smallest f32 estimate whose
 * window starts at s or later (breakpoints[0] = 0); buckets[1024]: window start at the
 * lower edge of each of 256 buckets per octave over [m/2, 8m). */
int hb_bias_windows(int p, double *mean, float *breakpoints, uint8_t *buckets);

int hb_eval_ranks(int device, const double *v, uint64_t n, uint32_t *rank);
int hb_eval_metrics(int device, const double *a, const double *b, uint64_t n, double top_frac,
                    double *out);

/* hb_eval_metrics_noalloc: hb_eval_metrics without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (tight serving or evaluation loops, harnesses that forbid hipMalloc).  a and b must
 * be device-accessible arrays (else HB_EINVAL); out is 4 host doubles, as above, and
 * the results are identical to hb_eval_metrics's.  Every intermediate buffer lives in
 * one fixed 32 MiB device array that the library holds on each device it runs on until
 * the process ends; n up to 2^19 fits, a larger n gives HB_ERANGE.  Runs on the
 * device's null stream and synchronises it.  Calls that overlap on one device share
 * that array and race: serialise them. */
int hb_eval_metrics_noalloc(int device, const double *a, const double *b, uint64_t n, double top_frac,
                            double *out);

/* ---- graph ingestion on the device ---------------------------------------------
 * hb_stage_mtx parses a MatrixMarket text ("%%MatrixMarket matrix coordinate
 * pattern|integer symmetric|skew-symmetric|general", square) on the device and stages
 * the resulting CSR as hb_stage_csr would: entry "i j" (1-based; further tokens
 * ignored) gives i-1 -> j-1, and j-1 -> i-1 too when the file is not "general" and
 * i != j; blank and '%' lines are skipped, a line with a single token is skipped,
 * duplicates are kept, every list is sorted ascending.  A malformed text returns
 * HB_EINVAL with info->err_code / err_line naming the first bad line.
 * Replaces: GraphCache::from_mtx_file (src/graph/cache/parser_mtx/mod.rs:34-162,
 * producers.rs:15-234, impl_sort.rs:10-51) up to the CSR it writes. */
typedef struct {
    uint64_t n, nnz;      /* size line                                              */
    uint64_t entries;     /* entry lines used                                       */
    uint64_t arcs;        /* arcs staged                                            */
    uint64_t err_line;    /* 1-based line of the first rejected line, 0 if none     */
    int32_t symmetric;    /* entries emit both ways (symmetric, skew-symmetric)     */
    int32_t err_code;     /* HB_MTX_*                                               */
    double ms_device;     /* device time of the line split, parse, sort and CSR     */
} hb_mtx_info;

#define HB_MTX_OK 0
#define HB_MTX_BAD_HEADER 1        /* first line does not start with %%MatrixMarket */
#define HB_MTX_MALFORMED_HEADER 2  /* fewer than 5 header tokens                    */
#define HB_MTX_UNSUPPORTED 3       /* not matrix coordinate pattern|integer ...     */
#define HB_MTX_NO_SIZE_LINE 4      /* text ends before the size line                */
#define HB_MTX_BAD_SIZE_LINE 5     /* size line lacks a number                      */
#define HB_MTX_NOT_SQUARE 6        /* rows != columns                               */
#define HB_MTX_BAD_ENTRY 7         /* an index is not a non-negative integer        */
#define HB_MTX_ZERO_INDEX 8        /* index 0 (MatrixMarket is 1-based)             */
#define HB_MTX_OUT_OF_BOUNDS 9     /* index above n                                 */

int hb_stage_mtx(hb_ctx *ctx, const char *text, uint64_t len, hb_mtx_info *info);

/* Builds and stages a CSR from an edge list on the device: arc src[k] -> dst[k] for
 * k < m (host or device arrays, ids < n), plus dst[k] -> src[k] when
 * HB_COO_SYMMETRIZE and src[k] != dst[k]; HB_COO_NO_LOOPS drops u -> u, HB_COO_DEDUP
 * keeps one copy of repeated arcs; lists sorted ascending.  No reference counterpart
 * (the synthetic generators' symmetrise / deduplicate step, on the device). */
#define HB_COO_SYMMETRIZE 0x1u
#define HB_COO_DEDUP 0x2u
#define HB_COO_NO_LOOPS 0x4u
int hb_stage_coo(hb_ctx *ctx, uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t m,
                 uint32_t flags);

/* Copies the staged CSR, in the caller's numbering, as 8-byte offsets[n+1] and
 * neighbours[arcs] (either may be NULL) -- the layout of the reference's index_<id> and
 * edges_<id> cache files (src/graph/cache/utils.rs). */
int hb_read_csr(hb_ctx *ctx, uint64_t *offsets, uint64_t *nbrs);

/* Number of nodes / arcs staged (0 if none). */
uint64_t hb_num_nodes(const hb_ctx *ctx);
uint64_t hb_num_arcs(const hb_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
