// hb_device.h -- device-side HyperLogLog arithmetic and shared kernel state for
// the MI355X HyperBall engine (gfx950, wave64).
//
// Register layout in HBM: node-major u8 registers, node u's 2^p registers at
// regs + u * 2^p, register j at byte j.  A "copy" of a sketch is held by
// G = 2^p / 16 lanes, each lane owning 16 consecutive registers in one uint4
// (p <= 10).  Every estimate is bit-identical to the reference's f32 value
// for the same registers (see estimate_copy below).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HB_WAVE 64

// ---------------------------------------------------------------------------
// Device state that survives across the launches of one run
// ---------------------------------------------------------------------------
struct HbIterState {
    uint32_t iter;       // iteration t being executed (1-based)
    uint32_t done;       // set by the finalize kernel; later launches return at once
    uint32_t executed;   // iterations executed when done
    uint32_t max_iters;  // 0: run to the reference's convergence rule
    uint32_t parity;     // registers read from regs[parity], written to regs[parity^1]
    uint32_t converged;
    double t, inv_t;     // t and 1/t as f64 (hyper_ball.rs:768-769,842-843)
    uint32_t skip_scan;  // most nodes changed last iteration: recompute all without the
                         // frontier scan (recomputing an unaffected node is a no-op)
    uint32_t push;       // few nodes changed last iteration (symmetric graph): build the
                         // list of nodes with a changed neighbour from the changed nodes
                         // instead of scanning every neighbour list
};

// kernel slots of one iteration: union bins 0..2, hub chunks, hub merge
#define HB_NSLOT 5
#define HB_SLOT_CHUNK 3
#define HB_SLOT_MERGE 4

// Per-iteration records live in a ring of HB_REC_RING entries (iteration t in slot
// t % HB_REC_RING, cleared when t begins); the host drains the finished ones at every
// check, which it makes at most 8 launches apart, so a run has no iteration cap.
#define HB_REC_RING 64

struct HbIterRec {  // per-iteration counters, per kernel slot
    unsigned long long inc[HB_NSLOT];    // nodes whose estimate increased
    unsigned long long arcs[HB_NSLOT];   // neighbour sketches gathered
    unsigned long long nodes[HB_NSLOT];  // nodes recomputed
    unsigned long long chg[HB_NSLOT];    // nodes whose registers changed
    unsigned long long gathered[HB_NSLOT];  // neighbour register blocks read (arcs counts ids scanned)
    unsigned long long nf_lo, nf_hi;     // N(t) * 2^40 as a 128-bit integer
    unsigned long long push;             // 1 if iteration t listed its nodes by push
};

struct HbBiasTable;  // HLL++ bias windows of one precision (below)

struct HbDev {
    const uint32_t *off;   // n+1
    const uint32_t *nbr;   // arcs
    // ids that are hashed: the caller's id of each node (null: the node's own index)
    // and of each arc's head (equal to nbr when the nodes keep the caller's order)
    const uint32_t *node_id;
    const uint32_t *orig;
    uint8_t *regs0, *regs1;
    float *est;            // estimate of each node's current sketch
    double *dist, *inv;    // accumulators
    uint32_t *bits0, *bits1;  // changed-node bitmaps (bit set: registers changed)
    HbIterState *st;
    HbIterRec *rec;        // ring: iteration t in rec[t % HB_REC_RING]
    const float *lc;       // lc[z] = (float)(m ln(m/z)), z = 1..m
    const HbBiasTable *bias;  // HLL++ bias windows for this p
    float *snaps;          // estimates after iterations 1..snap_k, iteration-major
    uint32_t snap_k;
    uint8_t *scratch;      // partial sketches of hub chunks
    uint32_t *chunk_flag;  // 1 if the chunk was recomputed this iteration
    // push iterations (symmetric graphs, p <= 10): nodes with a changed neighbour
    uint32_t *active;      // two bitmaps of (n + 31) / 32 words, by iteration parity (hb_active)
    uint32_t *dyn_list;    // their ids per degree bin, bin b at dyn_base[b]
    uint32_t *dyn_count;   // [3]
    uint32_t dyn_base[3];
    uint32_t small_max, mid_max, wave_max;  // degree bounds of bins 0..2 (hubs above)
    int sym;               // the CSR holds v -> u for every u -> v
    uint64_t arcs;         // arcs of the staged CSR
    uint32_t n;
    uint32_t n_linked;     // nodes with at least one neighbour (the scheduled ones)
    uint32_t p;
    float am;              // alpha_m * m * m, rounded to f32
    float five_m;          // 5m
    double lc_thr;         // HLL++ linear-counting threshold for p
    int frontier;
    int window;            // dense bin-0 iterations read in-window neighbours from LDS (chain layout)
    int rows;              // ... and, at p = 5 or 6, run one node per lane (k_union_rows)
};

// The active bitmap of iteration t.  There are two, by parity: iteration t marks its own
// in a push iteration and clears the other for iteration t + 1 (k_begin), so the marking
// does not wait for a clear.
__device__ __forceinline__ uint32_t *hb_active(const HbDev &d, uint32_t t)
{
    return d.active + (size_t)(t & 1u) * ((d.n + 31u) / 32u);
}

// ---------------------------------------------------------------------------
// SipHash-1-3 of an 8-byte little-endian message, key (0,0)
// (Aumasson & Bernstein; Rust's `usize: Hash` writes the value's 8 LE bytes)
// ---------------------------------------------------------------------------
// 64-bit rotate left by B (0 < B <= 32) as two 32-bit funnel shifts (v_alignbit_b32):
// the high word is the low 32 bits of (hi:lo) >> (32 - B), the low word of (lo:hi) >> (32 - B)
template <int B>
__device__ __forceinline__ uint64_t hb_rotl(uint64_t x)
{
    static_assert(B > 0 && B <= 32, "rotation");
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    if constexpr (B == 32) {
        return ((uint64_t)lo << 32) | hi;
    } else {
        const uint32_t nhi = __builtin_amdgcn_alignbit(hi, lo, 32 - B);
        const uint32_t nlo = __builtin_amdgcn_alignbit(lo, hi, 32 - B);
        return ((uint64_t)nhi << 32) | nlo;
    }
}

__device__ __forceinline__ uint64_t hb_siphash13_u64(uint64_t msg)
{
    uint64_t v0 = 0x736f6d6570736575ULL, v1 = 0x646f72616e646f6dULL;
    uint64_t v2 = 0x6c7967656e657261ULL, v3 = 0x7465646279746573ULL;
#define HB_SIPROUND                                                              \
    v0 += v1; v1 = hb_rotl<13>(v1); v1 ^= v0; v0 = hb_rotl<32>(v0);             \
    v2 += v3; v3 = hb_rotl<16>(v3); v3 ^= v2;                                    \
    v0 += v3; v3 = hb_rotl<21>(v3); v3 ^= v0;                                    \
    v2 += v1; v1 = hb_rotl<17>(v1); v1 ^= v2; v2 = hb_rotl<32>(v2);
    v3 ^= msg;
    HB_SIPROUND
    v0 ^= msg;
    const uint64_t b = 8ULL << 56;  // length byte, no tail bytes
    v3 ^= b;
    HB_SIPROUND
    v0 ^= b;
    v2 ^= 0xff;
    HB_SIPROUND HB_SIPROUND HB_SIPROUND
#undef HB_SIPROUND
    return v0 ^ v1 ^ v2 ^ v3;
}

// Register set by insert(x): index = top p bits of the hash, rank = 1 + leading
// zeros of (hash << p) with a sentinel bit at p-1 (rank <= 65-p < 64).
__device__ __forceinline__ void hb_index_rank(uint64_t x, uint32_t p, uint32_t *idx, uint32_t *rank)
{
    const uint64_t h = hb_siphash13_u64(x);
    *idx = (uint32_t)(h >> (64 - p));
    const uint64_t w = (h << p) | (1ULL << (p - 1));
    *rank = (uint32_t)__clzll((long long)w) + 1;
}

// ---------------------------------------------------------------------------
// Bytewise max of packed u8 registers with v_pk_max_u16: the even bytes and the
// odd bytes of a word are maxed as two 16-bit lanes each (the other byte of every
// lane masked to zero), then recombined.
// ---------------------------------------------------------------------------
typedef unsigned short hb_u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t hb_pkmax(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(hb_u16x2, a),
                                                                   __builtin_bit_cast(hb_u16x2, b)));
}

__device__ __forceinline__ uint32_t hb_bmax(uint32_t a, uint32_t b)
{
    return hb_pkmax(a & 0x00ff00ffu, b & 0x00ff00ffu) | hb_pkmax(a & 0xff00ff00u, b & 0xff00ff00u);
}

__device__ __forceinline__ uint4 hb_bmax4(uint4 a, uint4 b)
{
    return make_uint4(hb_bmax(a.x, b.x), hb_bmax(a.y, b.y), hb_bmax(a.z, b.z), hb_bmax(a.w, b.w));
}

// Running register-wise max over many sketches, kept as even-byte and odd-byte
// planes so that each gathered word costs two masks and two v_pk_max_u16.
struct HbAcc {
    uint32_t e[4], o[4];
};

__device__ __forceinline__ HbAcc hb_acc(uint4 x)
{
    HbAcc a;
    a.e[0] = x.x & 0x00ff00ffu; a.o[0] = x.x & 0xff00ff00u;
    a.e[1] = x.y & 0x00ff00ffu; a.o[1] = x.y & 0xff00ff00u;
    a.e[2] = x.z & 0x00ff00ffu; a.o[2] = x.z & 0xff00ff00u;
    a.e[3] = x.w & 0x00ff00ffu; a.o[3] = x.w & 0xff00ff00u;
    return a;
}

__device__ __forceinline__ void hb_acc_max(HbAcc &a, uint4 x)
{
    a.e[0] = hb_pkmax(a.e[0], x.x & 0x00ff00ffu); a.o[0] = hb_pkmax(a.o[0], x.x & 0xff00ff00u);
    a.e[1] = hb_pkmax(a.e[1], x.y & 0x00ff00ffu); a.o[1] = hb_pkmax(a.o[1], x.y & 0xff00ff00u);
    a.e[2] = hb_pkmax(a.e[2], x.z & 0x00ff00ffu); a.o[2] = hb_pkmax(a.o[2], x.z & 0xff00ff00u);
    a.e[3] = hb_pkmax(a.e[3], x.w & 0x00ff00ffu); a.o[3] = hb_pkmax(a.o[3], x.w & 0xff00ff00u);
}

__device__ __forceinline__ uint4 hb_acc_get(const HbAcc &a)
{
    return make_uint4(a.e[0] | a.o[0], a.e[1] | a.o[1], a.e[2] | a.o[2], a.e[3] | a.o[3]);
}

template <int W>
__device__ __forceinline__ uint4 hb_shfl_xor4(uint4 a, int mask)
{
    return make_uint4(__shfl_xor(a.x, mask, W), __shfl_xor(a.y, mask, W),
                      __shfl_xor(a.z, mask, W), __shfl_xor(a.w, mask, W));
}

// number of zero bytes (register values are < 128)
__device__ __forceinline__ uint32_t hb_zero_bytes(uint32_t x)
{
    const uint32_t nz = ((x | 0x80808080u) - 0x01010101u) & 0x80808080u;
    return 4u - (uint32_t)__popc(nz);
}

__device__ __forceinline__ uint32_t hb_max_byte(uint32_t x)
{
    x = hb_bmax(x, x >> 16);
    x = hb_bmax(x, x >> 8);
    return x & 0xffu;
}

// sum over the 4 bytes r of 2^(mx - r), each term clamped to 2^24
__device__ __forceinline__ uint32_t hb_scaled_terms(uint32_t x, uint32_t mx)
{
    uint32_t s = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t sh = mx - ((x >> (8 * b)) & 0xffu);
        s += 1u << (sh < 24u ? sh : 24u);
    }
    return s;
}

__device__ __forceinline__ float hb_pow2neg(uint32_t r) { return __uint_as_float((127u - r) << 23); }

// reductions over the G lanes of one sketch copy (G a power of two <= 64)
template <int G>
__device__ __forceinline__ uint32_t hb_copy_sum(uint32_t v)
{
#pragma unroll
    for (int s = 1; s < G; s <<= 1) v += __shfl_xor(v, s, G);
    return v;
}
template <int G>
__device__ __forceinline__ uint32_t hb_copy_max(uint32_t v)
{
#pragma unroll
    for (int s = 1; s < G; s <<= 1) v = max(v, (uint32_t)__shfl_xor(v, s, G));
    return v;
}
template <int G>
__device__ __forceinline__ uint32_t hb_copy_sum_sat24(uint32_t v)
{
    v = min(v, 1u << 24);
#pragma unroll
    for (int s = 1; s < G; s <<= 1) v = min(v + (uint32_t)__shfl_xor(v, s, G), 1u << 24);
    return v;
}
template <int G>
__device__ __forceinline__ uint32_t hb_copy_or(uint32_t v)
{
#pragma unroll
    for (int s = 1; s < G; s <<= 1) v |= __shfl_xor(v, s, G);
    return v;
}

// HLL++ EstimateBias (k = 6 nearest raw-estimate points, ties to the lower point,
// their biases summed in ascending order in f64 and divided by 6) as a lookup.  The
// window of 6 consecutive points the rule picks for an estimate e is a monotone step
// function of e, and e is an f32, so the host tabulates, per precision:
//   bp[s]   (s >= 1) the smallest f32 e whose window starts at point s or later;
//   mean[s] the window's mean bias, summed exactly as the rule sums it;
//   bucket[b] the window start at the lower edge of bucket b, 256 buckets per octave
//           over [m/2, 8m) (bucket = f32 bits >> 15, less bucket_base).
// The device reads the bucket's start and steps up while e >= bp[s + 1] (at most one
// or two steps: buckets are much narrower than the spacing of the points).
#define HB_BIAS_WINDOWS 195   // 200 points, windows of 6 consecutive points
#define HB_BIAS_BUCKETS 1024

struct HbBiasTable {
    double mean[HB_BIAS_WINDOWS + 1];
    float bp[HB_BIAS_WINDOWS + 1];  // bp[0] unused
    uint8_t bucket[HB_BIAS_BUCKETS];
    uint32_t bucket_base;           // (127 + p - 1) << 8: bucket 0 starts at m/2
    uint32_t pad;
};

__device__ __forceinline__ double hb_bias_lookup(const HbBiasTable *__restrict__ t, float e)
{
    const int32_t b = (int32_t)(__float_as_uint(e) >> 15) - (int32_t)t->bucket_base;
    uint32_t s = b < 0 ? 0u : t->bucket[b < HB_BIAS_BUCKETS ? b : HB_BIAS_BUCKETS - 1];
    while (s + 1 < HB_BIAS_WINDOWS && t->bp[s + 1] <= e) ++s;
    return t->mean[s];
}

// Lookup tables of the estimator for one p: lc[z] = (float)(m ln(m/z)), z = 1..m,
// and the HLL++ bias windows.  Kernels stage them in LDS.
struct HbTables {
    const float *lc;
    const HbBiasTable *bias;
    float am;       // alpha_m * m * m, rounded to f32
    float five_m;   // 5m
    double lc_thr;  // HLL++ linear-counting threshold for p
};

#define HB_TABLES_LDS(P, d, name)                                                      \
    __shared__ float name##_lc[(1 << (P)) + 1];                                        \
    __shared__ HbBiasTable name##_bias;                                                \
    for (uint32_t i_ = threadIdx.x; i_ <= (1u << (P)); i_ += blockDim.x) name##_lc[i_] = (d).lc[i_]; \
    for (uint32_t i_ = threadIdx.x; i_ < sizeof(HbBiasTable) / 8; i_ += blockDim.x)   \
        ((uint64_t *)&name##_bias)[i_] = ((const uint64_t *)(d).bias)[i_];           \
    __syncthreads();                                                                   \
    const HbTables name = {name##_lc, &name##_bias, (d).am, (d).five_m, (d).lc_thr};

// Estimate of the sketch held by a G-lane copy (each lane: 16 registers in acc).
// All lanes of the copy return the same value.
//   1. zeros > 0 and LC = m ln(m/zeros) <= threshold(p): return LC (HLL++).
//   2. raw = alpha m^2 / S with S = sum_j 2^-M_j summed in register order in f32.
//      S is computed exactly in integers: with mx = max_j M_j,
//      k = sum_j 2^(mx - M_j); when k < 2^24 every partial sum of the sequential
//      f32 loop is a multiple of 2^-mx below 2^(24-mx), hence exact, and
//      S = k * 2^-mx.  Otherwise the sequential loop is replayed.
//   3. raw <= 5m: subtract the HLL++ bias (k = 6 nearest neighbours, tabulated).
// f32 products and quotients are formed in f64 and rounded once, which gives the
// correctly rounded f32 result (53 >= 2*24 + 2).
template <int G>
__device__ __forceinline__ float hb_estimate_copy(const HbTables &tb, uint4 acc)
{
    uint32_t z = hb_zero_bytes(acc.x) + hb_zero_bytes(acc.y) + hb_zero_bytes(acc.z) +
                 hb_zero_bytes(acc.w);
    z = hb_copy_sum<G>(z);
    if (z > 0) {
        const float lc = tb.lc[z];
        if ((double)lc <= tb.lc_thr) return lc;
    }
    uint32_t mx = hb_max_byte(hb_bmax(hb_bmax(acc.x, acc.y), hb_bmax(acc.z, acc.w)));
    mx = hb_copy_max<G>(mx);
    uint32_t k = hb_scaled_terms(acc.x, mx) + hb_scaled_terms(acc.y, mx) +
                 hb_scaled_terms(acc.z, mx) + hb_scaled_terms(acc.w, mx);
    k = hb_copy_sum_sat24<G>(k);
    float S;
    if (k < (1u << 24)) {
        S = ldexpf((float)k, -(int)mx);
    } else {
        // replay the reference's sequential f32 loop over registers 0..m-1
        S = 0.0f;
        for (int l = 0; l < G; ++l) {
            const uint32_t w[4] = {(uint32_t)__shfl(acc.x, l, G), (uint32_t)__shfl(acc.y, l, G),
                                   (uint32_t)__shfl(acc.z, l, G), (uint32_t)__shfl(acc.w, l, G)};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int b = 0; b < 4; ++b) S = __fadd_rn(S, hb_pow2neg((w[q] >> (8 * b)) & 0xffu));
        }
    }
    float est = (float)((double)tb.am / (double)S);
    if (est <= tb.five_m) est = (float)((double)est - hb_bias_lookup(tb.bias, est));
    return est;
}

// The same estimate of a whole sketch held by one lane (W = 2^P/16 words, registers in
// order).  k sums at most 16 W terms of at most 2^24, so it cannot wrap, and k < 2^24
// exactly when the saturating sum of hb_estimate_copy is below 2^24: the same bits.
template <int W>
__device__ __forceinline__ float hb_estimate_row(const HbTables &tb, const uint4 *acc)
{
    uint32_t z = 0, mx = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const uint4 a = acc[q];
        z += hb_zero_bytes(a.x) + hb_zero_bytes(a.y) + hb_zero_bytes(a.z) + hb_zero_bytes(a.w);
        mx = max(mx, hb_max_byte(hb_bmax(hb_bmax(a.x, a.y), hb_bmax(a.z, a.w))));
    }
    if (z > 0) {
        const float lc = tb.lc[z];
        if ((double)lc <= tb.lc_thr) return lc;
    }
    uint32_t k = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const uint4 a = acc[q];
        k += hb_scaled_terms(a.x, mx) + hb_scaled_terms(a.y, mx) + hb_scaled_terms(a.z, mx) + hb_scaled_terms(a.w, mx);
    }
    float S;
    if (k < (1u << 24)) {
        S = ldexpf((float)k, -(int)mx);
    } else {
        S = 0.0f;  // the reference's sequential f32 loop over registers 0..m-1
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const uint4 a = acc[q];
            const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int b = 0; b < 4; ++b) S = __fadd_rn(S, hb_pow2neg((w[i] >> (8 * b)) & 0xffu));
        }
    }
    float est = (float)((double)tb.am / (double)S);
    if (est <= tb.five_m) est = (float)((double)est - hb_bias_lookup(tb.bias, est));
    return est;
}
