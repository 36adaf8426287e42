// pq_sa.hip -- suffix-array construction and FM-index build on the device.
//
// Suffix order (the reference's, sufsort.c:154-383 at one thread): suffixes
// run to the end of their read, end of read sorts first, ties by text
// position.
//
//  1. Bucket all N suffixes by a 2K+4-bit key: their first K codes (zero
//     padding past the read end) and min(length, K), K <= 14 (a 32-bit key).
//     The key is order-consistent, and a bucket whose length field is below K
//     holds identical complete suffixes, already in position order because the
//     LSD radix passes (from the identity permutation) are stable.
//  2. Find the bucket runs and split them into size classes with scans (no
//     global atomics): <= 16 and <= 64 elements are sorted by a group of lanes
//     with a bitonic network over registers, <= 2048 by one block in LDS.  The
//     comparator caches the first two 29-code chunk keys after the common
//     prefix and reads further chunks from the packed text only for suffixes
//     still tied (overlapping reads of one genome region).
//  3. Rare larger buckets are refined level by level: one chunk key per
//     element, global radix sort, and the tied sub-runs go back to step 2 at
//     the next chunk offset.
//
#include "pq_ctx.h"
#include "pq_radix.h"

#include <algorithm>
#include <cstring>
#include <chrono>

int pq_scan_u4(pq_ctx_t *ctx, const void *in, void *out, uint64_t count, void *total_host);

#define CHUNK 29

// ---- keys ------------------------------------------------------------------------

// Codes remaining from g to the end of its read.
__device__ __forceinline__ uint32_t suf_rem(const Geom &G, uint64_t g)
{
    return (uint32_t) pq_rem(G, g);
}

// The 29 codes of suffix g from offset `off`, bounded by the read end:
// (codes << 5) | min(codes remaining from g+off, 30).  Equal keys with a field
// below 30 mean the two suffixes end inside this chunk as the same string.
__device__ __forceinline__ uint64_t chunk_key(const uint64_t *t2, uint64_t g, uint32_t rem,
                                              uint32_t off)
{
    if (rem <= off)
        return 0;
    uint32_t f = rem - off;
    if (f > 30)
        f = 30;
    uint64_t chars = pq_window(t2, g + off) >> 6;
    if (f < CHUNK)
        chars &= ~((1ull << (2 * (CHUNK - f))) - 1ull);
    return (chars << 5) | f;
}

// Bucket key of the first K codes (K <= 14): (codes << 4) | min(rem, K).  Keys
// with a field below K are complete suffixes; equal ones are equal strings.
__device__ __forceinline__ uint32_t bucket_key(const Geom &G, const uint64_t *t2, uint64_t g,
                                               unsigned K)
{
    uint32_t rem = suf_rem(G, g);
    uint32_t f = rem < K ? rem : K;
    uint32_t chars = (uint32_t) (pq_window(t2, g) >> (64 - 2 * K));
    if (f < K)
        chars &= ~((1u << (2 * (K - f))) - 1u);
    return (chars << 4) | f;
}

// BWT symbol of the row holding suffix g: the code before g, or '$' (0) at a
// read start.  Carried through the sort and written next to the row it
// finalises.
__device__ __forceinline__ unsigned bwt_sym(const Geom &G, const uint64_t *t2, uint64_t g)
{
    return pq_is_start(G, g) ? 0u : pq_code(t2, g - 1) + 1u;
}

// With 64-bit rows a suffix-array value carries its row's BWT symbol in its top
// three bits (position | symbol << 61) from the bucket pass until a sort writes
// the row or k_bwt_fill finalises it: the run sorts take the symbol from there,
// and the rows no sort touches get theirs without a text read.  32-bit rows have
// no room; their symbols come from the text.
#define SA_SYM_SHIFT 61
template <typename R> __device__ __forceinline__ uint64_t sa_pos(R v)
{
    if constexpr (sizeof(R) == 8)
        return (uint64_t) v & ((1ull << SA_SYM_SHIFT) - 1ull);
    else
        return (uint64_t) v;
}
template <typename R> __device__ __forceinline__ R sa_pack(uint64_t g, unsigned sym)
{
    if constexpr (sizeof(R) == 8)
        return (R) (g | ((uint64_t) sym << SA_SYM_SHIFT));
    else
        return (R) g;
}

// The (bucket key, position) pairs of all N suffixes in position order, made
// on the fly by the first radix pass of a build in one slice (no key array is
// written and read back).
template <typename R> struct BucketSrc {
    Geom G;
    const uint64_t *t2;
    unsigned K;
    __device__ __forceinline__ uint32_t key(uint64_t g) const { return bucket_key(G, t2, g, K); }
    __device__ __forceinline__ void load(uint64_t g, uint32_t &k, R &v) const
    {
        k = bucket_key(G, t2, g, K);
        v = sa_pack<R>(g, sizeof(R) == 8 ? bwt_sym(G, t2, g) : 0u);
    }
};

// ---- slices of the suffix set ---------------------------------------------------------
//
// For large texts the suffixes are sorted a slice at a time: a slice is a range
// of bucket keys, so it is a contiguous range of the suffix array, and its
// scratch is bounded by the slice size instead of N.

#define SLICE_BINS 4096      // histogram over the top 12 key bits
#define SLICE_TILE 8192      // positions per block of the slice kernels (32 per thread)

// Bucket keys of the text positions 32 j .. 32 j + 31 below N, in order:
// keys[i] for position 32 j + i (i < returned count).  The codes come from two
// words loaded once and the codes left in the read are carried from position
// to position: one read lookup per 32 positions instead of a window load and a
// read lookup per position, which made the slice passes compute-bound.
template <bool SYM = false>
__device__ __forceinline__ uint32_t bucket_keys32(const Geom &G, const uint64_t *t2, uint64_t N,
                                                  unsigned K, uint64_t j, uint32_t keys[32],
                                                  uint8_t *syms = nullptr)
{
    const uint64_t g0 = j << 5;
    if (g0 >= N)
        return 0;
    const uint64_t w0 = t2[j], w1 = t2[j + 1];   // slack words follow the text
    uint64_t r = pq_read_of(G, g0);
    uint32_t rlen = pq_read_len(G, r);
    uint32_t rem = (uint32_t) (pq_read_start(G, r) + rlen - g0);
    const uint32_t n = N - g0 < 32 ? (uint32_t) (N - g0) : 32u;
    // with SYM, syms[i] = BWT symbol of the row of position 32 j + i: '$' (0) at
    // a read start, else the code before it + 1
    uint32_t prev = 0;
    if constexpr (SYM)
        prev = j ? (uint32_t) (t2[j - 1] & 3u) : 0u;
#pragma unroll
    for (uint32_t i = 0; i < 32; i++) {
        if (i < n) {
            while (rem == 0) {
                rlen = pq_read_len(G, ++r);
                rem = rlen;
            }
            if constexpr (SYM)
                syms[i] = (uint8_t) (rem == rlen ? 0u : prev + 1u);
            const unsigned sh = 2 * i;
            const uint64_t win = (w0 << sh) | ((w1 >> 1) >> (63 - sh));
            const uint32_t f = rem < K ? rem : K;
            uint32_t chars = (uint32_t) (win >> (64 - 2 * K));
            if (f < K)
                chars &= ~((1u << (2 * (K - f))) - 1u);
            keys[i] = (chars << 4) | f;
            if constexpr (SYM)
                prev = (uint32_t) (w0 >> (62 - 2 * i)) & 3u;
            rem--;
        }
    }
    return n;
}

__global__ void __launch_bounds__(256) k_key_hist(Geom G, const uint64_t *t2, uint64_t N,
                                                  unsigned K, unsigned shift,
                                                  unsigned long long *hist)
{
    __shared__ uint32_t h[SLICE_BINS];
    for (uint32_t i = threadIdx.x; i < SLICE_BINS; i += blockDim.x)
        h[i] = 0;
    __syncthreads();
    PQ_GRID_STRIDE_FOR(j, (N + 31) >> 5)
    {
        uint32_t keys[32];
        const uint32_t n = bucket_keys32(G, t2, N, K, j, keys);
#pragma unroll
        for (uint32_t i = 0; i < 32; i++)
            if (i < n)
                atomicAdd(&h[keys[i] >> shift], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < SLICE_BINS; i += blockDim.x)
        if (h[i])
            atomicAdd(&hist[i], (unsigned long long) h[i]);
}

// Block sum of one value per thread (256 threads); the result in every thread.
__device__ __forceinline__ uint32_t block_sum256(uint32_t c, uint32_t *ws)
{
    for (int d = 32; d > 0; d >>= 1)
        c += __shfl_xor(c, d, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        ws[threadIdx.x >> 6] = c;
    __syncthreads();
    return ws[0] + ws[1] + ws[2] + ws[3];
}

// Suffixes of the slice [lo, hi) of keys per tile of SLICE_TILE positions.
__global__ void __launch_bounds__(256) k_slice_count(Geom G, const uint64_t *t2, uint64_t N,
                                                     unsigned K, uint64_t lo, uint64_t hi,
                                                     uint32_t *bcnt)
{
    __shared__ uint32_t ws[4];
    uint32_t keys[32];
    const uint64_t j = (uint64_t) blockIdx.x * (SLICE_TILE / 32) + threadIdx.x;
    const uint32_t n = bucket_keys32(G, t2, N, K, j, keys);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t i = 0; i < 32; i++)
        c += (i < n && keys[i] >= lo && keys[i] < hi) ? 1u : 0u;
    c = block_sum256(c, ws);
    if (threadIdx.x == 0)
        bcnt[blockIdx.x] = c;
}

// Slice bounds for k_slice_count_all, clamped to 32 bits: b[s] = first key of
// slice s; past the last slice all ones, which no key reaches (a key's low four
// bits, its length field, are at most 14).
#define SLICE_MAX 16
struct SliceBounds {
    uint32_t b[SLICE_MAX + 1];
};
// First suffix-array row of every slice.
struct SliceStarts {
    uint64_t start[SLICE_MAX];
};

// Suffixes of every slice per tile of SLICE_TILE positions, in one pass over the
// text: bcnt[s * ntiles + tile] for slice s = [b[s], b[s + 1]) of keys, from the
// counts of keys at or above each bound (bound 0 is key 0).  NS >= the slice count.
template <int NS>
__global__ void __launch_bounds__(256) k_slice_count_all(Geom G, const uint64_t *t2, uint64_t N,
                                                         unsigned K, SliceBounds sb, uint32_t nsl,
                                                         uint64_t ntiles, uint32_t *bcnt)
{
    __shared__ uint32_t part[4][NS];
    uint32_t keys[32];
    const uint64_t j = (uint64_t) blockIdx.x * (SLICE_TILE / 32) + threadIdx.x;
    const uint32_t n = bucket_keys32(G, t2, N, K, j, keys);
    uint32_t ge[NS];
    ge[0] = n;
#pragma unroll
    for (int s = 1; s < NS; s++)
        ge[s] = 0;
#pragma unroll
    for (uint32_t i = 0; i < 32; i++)
        if (i < n)
#pragma unroll
            for (int s = 1; s < NS; s++)
                ge[s] += keys[i] >= sb.b[s] ? 1u : 0u;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        uint32_t c = ge[s];
        for (int d = 32; d > 0; d >>= 1)
            c += __shfl_xor(c, d, 64);
        if (lane == 0)
            part[wid][s] = c;
    }
    __syncthreads();
    const uint32_t s = threadIdx.x;
    if (s < nsl) {
        const uint32_t a = part[0][s] + part[1][s] + part[2][s] + part[3][s];
        const uint32_t b = s + 1 < (uint32_t) NS
                               ? part[0][s + 1] + part[1][s + 1] + part[2][s + 1] + part[3][s + 1]
                               : 0u;
        bcnt[(uint64_t) s * ntiles + blockIdx.x] = a - b;
    }
}

// Keys and positions of the slice's suffixes, in position order (the radix
// passes that follow are stable, and runs of identical complete suffixes rely
// on position order).  A lane makes the keys of 32 consecutive positions and
// parks them in LDS; its wave then takes its 2048 positions 64 at a time, so the
// slice members of each round are written as one coalesced run.
template <typename R>
__global__ void __launch_bounds__(256) k_slice_emit(Geom G, const uint64_t *t2, uint64_t N,
                                                    unsigned K, uint64_t lo, uint64_t hi,
                                                    const uint32_t *boff, uint32_t *BK, R *V)
{
    __shared__ uint32_t sk[4][64 * 33];   // a lane's 32 keys 33 words apart: no bank conflicts
    __shared__ uint32_t ssy[4][64 * 5];   // their BWT symbols, 4 bits each (64-bit rows)
    __shared__ uint32_t ws[4];
    constexpr bool SYM = sizeof(R) == 8;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t keys[32];
    uint8_t syms[32] = {};
    const uint64_t j = (uint64_t) blockIdx.x * (SLICE_TILE / 32) + threadIdx.x;
    const uint32_t n = bucket_keys32<SYM>(G, t2, N, K, j, keys, syms);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t i = 0; i < 32; i++) {
        const bool in = i < n && keys[i] >= lo && keys[i] < hi;
        c += in ? 1u : 0u;
        // all ones marks a non-member: no key has 15 in its length field
        sk[wid][33 * lane + i] = in ? keys[i] : 0xFFFFFFFFu;
    }
    if constexpr (SYM) {
        // eight symbols per word, five words per lane (stride 5: no bank conflicts)
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            uint32_t x = 0;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++)
                x |= (uint32_t) syms[8 * q + i] << (4 * i);
            ssy[wid][5 * lane + q] = x;
        }
    }
    for (int d = 32; d > 0; d >>= 1)
        c += __shfl_xor(c, d, 64);
    if (lane == 0)
        ws[wid] = c;
    __syncthreads();
    uint32_t base = boff[blockIdx.x];
    for (int w = 0; w < wid; w++)
        base += ws[w];
    const uint64_t g0 = (uint64_t) blockIdx.x * SLICE_TILE + (uint64_t) wid * 2048;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t q = 0; q < 32; q++) {
        const uint32_t p = q * 64 + lane;
        const uint32_t k = sk[wid][33 * (p >> 5) + (p & 31)];
        const bool in = k != 0xFFFFFFFFu;
        const uint64_t m = __ballot(in);
        if (in) {
            const uint32_t at = base + (uint32_t) __popcll(m & lt);
            BK[at] = k;
            unsigned sym = 0;
            if constexpr (SYM)
                sym = (ssy[wid][5 * (p >> 5) + ((p & 31) >> 3)] >> (4 * (p & 7))) & 15u;
            V[at] = sa_pack<R>(g0 + p, sym);
        }
        base += (uint32_t) __popcll(m);
    }
}

// Every slice's (key, position) pairs at once, each into its slice's place
// (start[s] + the slice's exclusive prefix of tile counts soff[s * ntiles +
// tile] + the rank within the tile), in position order within each slice: the
// bucket keys are made once instead of once per slice.  As k_slice_emit
// otherwise; NS >= the slice count.
template <typename R, int NS>
__global__ void __launch_bounds__(256)
    k_slice_emit_all(Geom G, const uint64_t *t2, uint64_t N, unsigned K, SliceBounds sb,
                     SliceStarts ss, uint32_t nsl, uint64_t ntiles, const uint32_t *soff,
                     uint32_t *BK, R *V)
{
    __shared__ uint32_t sk[4][64 * 33];
    __shared__ uint32_t ssy[4][64 * 5];
    __shared__ uint32_t wsc[4][NS];
    constexpr bool SYM = sizeof(R) == 8;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t keys[32];
    uint8_t syms[32] = {};
    const uint64_t j = (uint64_t) blockIdx.x * (SLICE_TILE / 32) + threadIdx.x;
    const uint32_t n = bucket_keys32<SYM>(G, t2, N, K, j, keys, syms);
    // the slice of a key: how many bounds after the first it reaches
    auto slice_of = [&](uint32_t k) -> uint32_t {
        uint32_t sl = 0;
#pragma unroll
        for (int t = 1; t < NS; t++)
            sl += k >= sb.b[t] ? 1u : 0u;
        return sl;
    };
    uint32_t c[NS];
#pragma unroll
    for (int t = 0; t < NS; t++)
        c[t] = 0;
#pragma unroll
    for (uint32_t i = 0; i < 32; i++) {
        const uint32_t sl = slice_of(keys[i]);
#pragma unroll
        for (int t = 0; t < NS; t++)
            c[t] += (i < n && sl == (uint32_t) t) ? 1u : 0u;
        sk[wid][33 * lane + i] = i < n ? keys[i] : 0xFFFFFFFFu;
    }
    if constexpr (SYM) {
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            uint32_t x = 0;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++)
                x |= (uint32_t) syms[8 * q + i] << (4 * i);
            ssy[wid][5 * lane + q] = x;
        }
    }
#pragma unroll
    for (int t = 0; t < NS; t++) {
        uint32_t x = c[t];
        for (int d = 32; d > 0; d >>= 1)
            x += __shfl_xor(x, d, 64);
        if (lane == 0)
            wsc[wid][t] = x;
    }
    __syncthreads();
    uint64_t base[NS];
#pragma unroll
    for (int t = 0; t < NS; t++) {
        base[t] = 0;
        if ((uint32_t) t < nsl) {
            base[t] = ss.start[t] + soff[(uint64_t) t * ntiles + blockIdx.x];
            for (int w = 0; w < wid; w++)
                base[t] += wsc[w][t];
        }
    }
    const uint64_t g0 = (uint64_t) blockIdx.x * SLICE_TILE + (uint64_t) wid * 2048;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t q = 0; q < 32; q++) {
        const uint32_t p = q * 64 + lane;
        const uint32_t k = sk[wid][33 * (p >> 5) + (p & 31)];
        const bool valid = k != 0xFFFFFFFFu;   // no key has 15 in its length field
        const uint32_t sl = slice_of(k);
        unsigned sym = 0;
        if constexpr (SYM)
            sym = (ssy[wid][5 * (p >> 5) + ((p & 31) >> 3)] >> (4 * (p & 7))) & 15u;
#pragma unroll
        for (int t = 0; t < NS; t++) {
            if ((uint32_t) t < nsl) {
                const bool mine = valid && sl == (uint32_t) t;
                const uint64_t m = __ballot(mine);
                if (mine) {
                    const uint64_t at = base[t] + (uint64_t) __popcll(m & lt);
                    BK[at] = k;
                    V[at] = sa_pack<R>(g0 + p, sym);
                }
                base[t] += (uint64_t) __popcll(m);
            }
        }
    }
}

// One suffix of a run being sorted: the codes left after the run's common
// prefix of `base` codes packed in W words (32 codes each, most significant
// first, zero past the end), then tie = len << 44 | position << 3 | sym: that
// count of codes (below 2^20), the text position (below 2^41) and the BWT symbol
// of the row the suffix lands in.  Words then tie order suffixes by codes, then
// length, then position (positions are unique: the symbol never decides); a
// padding element is all ones.  Twelve 32-bit registers for W = 5, which every
// step of a sorting network exchanges and selects.  All of it is loaded up front
// with independent loads, so comparisons never wait on memory.
template <int W, typename R> struct SufW {
    uint64_t w[W];
    uint64_t tie;
    __device__ __forceinline__ R g() const { return (R) ((tie >> 3) & ((1ull << 41) - 1ull)); }
    __device__ __forceinline__ uint32_t len() const { return (uint32_t) (tie >> 44); }
    __device__ __forceinline__ unsigned sym() const { return (unsigned) (tie & 7u); }
};


#define BWT_UNSET 0xFFu

template <int W, typename R>
__device__ __forceinline__ void suf_load(const Geom &G, const uint64_t *t2, R v, uint32_t base,
                                         SufW<W, R> &s)
{
    const uint64_t g = sa_pos<R>(v);   // v: a suffix-array value of the slice being sorted
    const uint32_t rem = suf_rem(G, g);
    const uint32_t len = rem > base ? rem - base : 0;
    // The W + 1 text words the windows come from and the word holding the code
    // before g, all loaded unconditionally (word indices clamped to the packed
    // text's slack words), so that they are in flight together: a load under a
    // length branch is waited for before the next one is issued.
    const uint64_t wmax = (G.N >> 5) + 3;
    const uint64_t p0 = (uint64_t) g + base, w0 = p0 >> 5;
    uint64_t raw[W + 1];
#pragma unroll
    for (int i = 0; i <= W; i++)
        raw[i] = t2[w0 + i < wmax ? w0 + i : wmax];
    unsigned sym;
    if constexpr (sizeof(R) == 8) {
        sym = (unsigned) ((uint64_t) v >> SA_SYM_SHIFT);
    } else {
        const uint64_t gp = g ? g - 1 : 0;
        const uint64_t prev = t2[gp >> 5];
        sym = pq_is_start(G, g) ? 0u : (unsigned) ((prev >> (62 - 2 * (gp & 31))) & 3u) + 1u;
    }
    s.tie = ((uint64_t) len << 44) | (g << 3) | sym;
    const unsigned sh = (unsigned) (p0 & 31) * 2;
#pragma unroll
    for (int i = 0; i < W; i++) {
        // window i: 32 codes from p0 + 32 i, zero past the suffix's end
        uint64_t x = (raw[i] << sh) | ((raw[i + 1] >> 1) >> (63 - sh));
        const uint32_t lo = 32u * (uint32_t) i;
        if (lo >= len)
            x = 0;
        else if (len - lo < 32)
            x &= ~0ull << (64 - 2 * (len - lo));
        s.w[i] = x;
    }
}

// A padding element: all-ones words, length and position, so that it sorts after
// every suffix.
template <int W, typename R> __device__ __forceinline__ void suf_none(SufW<W, R> &s)
{
    s.tie = ~0ull;
#pragma unroll
    for (int i = 0; i < W; i++)
        s.w[i] = ~0ull;
}

// Sampled keys of the learned index written by the run sorts as they place their
// rows: the key of a row that is a multiple of 1 << shift is its suffix's first D
// codes, zero past the end of its read, | (the suffix has at least D codes) --
// suffix_key, the key the learned index samples.  First-level runs (sharing
// exactly their bucket's K codes) make it from the bucket's codes and the cached
// words; deeper levels (from_text) read it from the text.  Rows no sort places
// get theirs from k_bwt_fill.
struct KeyOut {
    u128 *keys;              // nullptr: no keys
    const uint32_t *BK;      // bucket keys of the slice's rows
    uint64_t row0;           // first row of the slice
    uint32_t shift, D, K;
    uint32_t from_text;
};

template <int W, typename R>
__device__ __forceinline__ void emit_key(const KeyOut &ko, const Geom &G, const uint64_t *t2,
                                         uint32_t row, uint32_t bk, const SufW<W, R> &s)
{
    const uint64_t gr = ko.row0 + row;
    if (gr & ((1ull << ko.shift) - 1ull))
        return;
    if (ko.from_text) {
        ko.keys[gr >> ko.shift] = suffix_key(G, t2, (uint64_t) s.g(), ko.D);
        return;
    }
    const uint32_t rem = ko.K + s.len();   // every suffix of such a run has >= K codes
    uint64_t w1 = 0;
    if constexpr (W > 1)
        w1 = s.w[1];
    const u128 w = ((u128) s.w[0] << 64) | (u128) w1;
    u128 key = ((u128) (bk >> 4) << (128 - 2 * ko.K)) | (w >> (2 * ko.K));
    key &= top_mask(ko.D);
    ko.keys[gr >> ko.shift] = key | (u128) (rem >= ko.D ? 1u : 0u);
}

// Canonical order of two suffixes sharing their first `base` codes: string
// order with the end of the read first, then position.  The cached words are zero
// past a suffix's end, so comparing the words, then the lengths, then the
// positions gives that order (a proper prefix is smaller or equal word by word,
// then shorter) with no branch; padding elements (suf_none) compare last.
template <int W, typename R>
__device__ __forceinline__ bool suf_less(const SufW<W, R> &a, const SufW<W, R> &b)
{
    bool lt = false, eq = true;
#pragma unroll
    for (int i = 0; i < W; i++) {
        lt = lt | (eq & (a.w[i] < b.w[i]));
        eq = eq & (a.w[i] == b.w[i]);
    }
    return lt | (eq & (a.tie < b.tie));
}

template <int J, int W, typename R>
__device__ __forceinline__ SufW<W, R> suf_xor(const SufW<W, R> &s)
{
    SufW<W, R> o;
    o.tie = xor_lane64<J>(s.tie);
#pragma unroll
    for (int i = 0; i < W; i++)
        o.w[i] = xor_lane64<J>(s.w[i]);
    return o;
}

// Bitonic network over the lanes of a group: stage K (merge width), step J.
template <int K, int J, int W, typename R>
__device__ __forceinline__ void bitonic_steps(SufW<W, R> &s, int li)
{
    SufW<W, R> p = suf_xor<J, W, R>(s);
    const bool up = (li & K) == 0;
    const bool lower = (li & J) == 0;
    const bool pless = suf_less<W, R>(p, s);
    const bool swap = (lower == up) ? pless : (!pless && p.tie != s.tie);
    if (swap)
        s = p;
    if constexpr (J > 1)
        bitonic_steps<K, J / 2, W, R>(s, li);
}

template <int G, int K, int W, typename R>
__device__ __forceinline__ void bitonic_stages(SufW<W, R> &s, int li)
{
    bitonic_steps<K, K / 2, W, R>(s, li);
    if constexpr (K < G)
        bitonic_stages<G, 2 * K, W, R>(s, li);
}

// ---- run lists -----------------------------------------------------------------------

#define NCLS 7              // run size classes, see seg_class

struct SegLists {
    uint2 *items[NCLS];
    uint32_t *count;         // NCLS counters
    uint32_t cap[NCLS];
    uint32_t *overflow;
    uint32_t lcap;           // largest run of the LDS class (0: words too long)
};

#define RUN_SMALL 16
#define RUN_HALF 32
#define RUN_WAVE 64
#define RUN_WAVE2 128       // two suffixes per lane of one wave
#define RUN_WAVE4 256       // four per lane
#define CLS_BIG (NCLS - 1)
#define CLS_DONE 255

// 0: <= 16 elements, 1: <= 32, 2: <= 64 (lanes of one wave), 3: <= 128 and
// 4: <= 256 (registers of one wave), 5: <= lcap (LDS, or the registers of one
// wave when lcap is 512), 6: larger.
// lcap 0 means the remaining suffix length exceeds the cached words: everything
// goes to the chunk-level path.
__device__ __forceinline__ int seg_class(uint32_t size, uint32_t lcap)
{
    if (lcap == 0)
        return CLS_BIG;
    return size <= RUN_SMALL   ? 0
           : size <= RUN_HALF  ? 1
           : size <= RUN_WAVE  ? 2
           : size <= RUN_WAVE2 ? 3
           : size <= RUN_WAVE4 ? 4
           : size <= lcap      ? 5
                               : CLS_BIG;
}

// Wave-aggregated append (used only for the runs split off oversized runs);
// every lane of the wave must call it.
__device__ __forceinline__ void push_seg(const SegLists &S, bool valid, uint32_t start,
                                         uint32_t size)
{
    int lane = threadIdx.x & 63;
    int cls = seg_class(size, S.lcap);
    uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < NCLS; c++) {
        bool mine = valid && cls == c;
        uint64_t m = __ballot(mine);
        if (m) {
            int leader = __ffsll((long long) m) - 1;
            uint32_t base = 0;
            if (lane == leader)
                base = atomicAdd(&S.count[c], (uint32_t) __popcll(m));
            base = __shfl(base, leader, 64);
            if (mine) {
                uint32_t at = base + (uint32_t) __popcll(m & lt);
                if (at < S.cap[c])
                    S.items[c][at] = make_uint2(start, size);
                else
                    atomicOr(S.overflow, 1u);
            }
        }
    }
}

// ---- sorts of one run ---------------------------------------------------------------

// One group of G lanes per run (G = 16 or 64); bitonic network over lanes.
template <int G, int W, typename R>
__global__ void __launch_bounds__(256) k_sort_groups(Geom Gm, const uint64_t *t2, R *sa,
                                                     uint8_t *bwt8, uint32_t base,
                                                     const uint2 *segs, uint32_t nseg, KeyOut ko)
{
    uint64_t tid = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t sid = tid / G;
    int li = (int) (threadIdx.x % G);
    bool have = sid < nseg;
    uint2 sg = have ? segs[sid] : make_uint2(0, 0);
    bool e = have && (uint32_t) li < sg.y;
    // the run's bucket key (for the sampled keys) loaded with the suffix, not
    // after the sort
    const uint32_t bk = e && ko.keys && !ko.from_text ? ko.BK[sg.x] : 0u;
    SufW<W, R> s;
    if (e)
        suf_load<W, R>(Gm, t2, sa[sg.x + li], base, s);
    else
        suf_none<W, R>(s);
    bitonic_stages<G, 2, W, R>(s, li);
    if (e) {
        sa[sg.x + li] = s.g();
        bwt8[sg.x + li] = (uint8_t) s.sym();
        if (ko.keys)
            emit_key<W, R>(ko, Gm, t2, sg.x + li, bk, s);
    }
}

// Bitonic network over M x 64 suffixes held M per lane by one wave: element
// e = m * 64 + lane.  Steps J < 64 pair lanes (DPP / swizzle / permute), steps
// J >= 64 pair items m and m ^ (J / 64) of the same lane.
template <int M, int K, int J, int W, typename R>
__device__ __forceinline__ void wave_steps(SufW<W, R> (&s)[M], int lane)
{
    if constexpr (J >= 64) {
        constexpr int JM = J / 64;
#pragma unroll
        for (int m = 0; m < M; m++) {
            if ((m & JM) == 0) {
                // the lower item keeps the smaller of the pair when the merge runs up
                const bool up = ((m * 64) & K) == 0;
                const bool bless = suf_less<W, R>(s[m + JM], s[m]);
                if (up ? bless : (!bless && s[m + JM].tie != s[m].tie)) {
                    const SufW<W, R> x = s[m];
                    s[m] = s[m + JM];
                    s[m + JM] = x;
                }
            }
        }
    } else {
        const bool lower = (lane & J) == 0;
#pragma unroll
        for (int m = 0; m < M; m++) {
            const bool up = ((m * 64 + lane) & K) == 0;
            const SufW<W, R> p = suf_xor<J, W, R>(s[m]);
            const bool pless = suf_less<W, R>(p, s[m]);
            const bool swap = (lower == up) ? pless : (!pless && p.tie != s[m].tie);
            if (swap)
                s[m] = p;
        }
    }
    if constexpr (J > 1)
        wave_steps<M, K, J / 2, W, R>(s, lane);
}

template <int M, int K, int W, typename R>
__device__ __forceinline__ void wave_stages(SufW<W, R> (&s)[M], int lane)
{
    wave_steps<M, K, K / 2, W, R>(s, lane);
    if constexpr (K < 64 * M)
        wave_stages<M, 2 * K, W, R>(s, lane);
}

// One wave per run of up to M x 64 suffixes (4 runs per block), sorted in
// registers: all of a lane's suffix words are loaded up front, and no step
// waits on LDS or a block barrier.
template <int M, int W, typename R>
__global__ void __launch_bounds__(256) k_sort_wave(Geom Gm, const uint64_t *t2, R *sa,
                                                   uint8_t *bwt8, uint32_t base,
                                                   const uint2 *segs, uint32_t nseg, KeyOut ko)
{
    const int lane = threadIdx.x & 63;
    const uint64_t sid = ((uint64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (sid >= nseg)
        return;   // whole waves: sid is uniform over a wave
    const uint2 sg = segs[sid];
    SufW<W, R> s[M];
    // every item loaded (the run's last one past its end: sg.y >= 1), then the
    // padding marked, so that no load waits behind a branch
    R gs[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const uint32_t e = (uint32_t) (m * 64 + lane);
        gs[m] = sa[sg.x + (e < sg.y ? e : sg.y - 1)];
    }
#pragma unroll
    for (int m = 0; m < M; m++)
        suf_load<W, R>(Gm, t2, gs[m], base, s[m]);
#pragma unroll
    for (int m = 0; m < M; m++)
        if ((uint32_t) (m * 64 + lane) >= sg.y)
            suf_none<W, R>(s[m]);
    const uint32_t bk = ko.keys && !ko.from_text ? ko.BK[sg.x] : 0u;
    wave_stages<M, 2, W, R>(s, lane);
#pragma unroll
    for (int m = 0; m < M; m++) {
        const uint32_t e = (uint32_t) (m * 64 + lane);
        if (e < sg.y) {
            sa[sg.x + e] = s[m].g();
            bwt8[sg.x + e] = (uint8_t) s[m].sym();
            if (ko.keys)
                emit_key<W, R>(ko, Gm, t2, sg.x + e, bk, s[m]);
        }
    }
}

// Largest run one 256-thread block sorts in LDS for W words per suffix.
__host__ __device__ constexpr uint32_t lds_cap(int W)
{
    return W <= 2 ? 2048u : W <= 4 ? 1024u : 512u;
}

// One block per run of 65..lds_cap(W) elements: bitonic sort in LDS.
template <int W, typename R>
__global__ void __launch_bounds__(256) k_sort_lds(Geom Gm, const uint64_t *t2, R *sa,
                                                  uint8_t *bwt8, uint32_t base, const uint2 *segs,
                                                  KeyOut ko)
{
    constexpr uint32_t CAP = lds_cap(W);
    __shared__ SufW<W, R> sh[CAP];
    uint2 sg = segs[blockIdx.x];
    uint32_t n = sg.y, P = 1;
    while (P < n)
        P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        SufW<W, R> s;
        if (i < n)
            suf_load<W, R>(Gm, t2, sa[sg.x + i], base, s);
        else
            suf_none<W, R>(s);
        sh[i] = s;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = threadIdx.x; p < P / 2; p += blockDim.x) {
                uint32_t i = 2 * j * (p / j) + (p % j), ip = i + j;
                bool up = (i & k) == 0;
                SufW<W, R> x = sh[i], y = sh[ip];
                if (suf_less<W, R>(y, x) == up) {
                    sh[i] = y;
                    sh[ip] = x;
                }
            }
            __syncthreads();
        }
    }
    const uint32_t bk = ko.keys && !ko.from_text ? ko.BK[sg.x] : 0u;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        sa[sg.x + i] = sh[i].g();
        bwt8[sg.x + i] = (uint8_t) sh[i].sym();
        if (ko.keys)
            emit_key<W, R>(ko, Gm, t2, sg.x + i, bk, sh[i]);
    }
}

// ---- runs of equal keys ---------------------------------------------------------------

#define RUN_TILE 4096        // keys per block of the run-start kernels (16 x 256)

// Run heads (keys that differ from the key before them) per tile of RUN_TILE keys.
template <typename KT>
__global__ void __launch_bounds__(256) k_run_counts(const KT *K, uint64_t n, uint32_t *tcnt)
{
    __shared__ uint32_t ws[4];
    const uint64_t tile = (uint64_t) blockIdx.x * RUN_TILE;
    // keys at clamped indices, loaded without a bounds branch (loads under one are
    // waited for in turn)
    KT a[RUN_TILE / 256], b[RUN_TILE / 256];
#pragma unroll
    for (int r = 0; r < RUN_TILE / 256; r++) {
        const uint64_t i = tile + (uint64_t) r * 256 + threadIdx.x;
        const uint64_t ic = i < n ? i : n - 1;
        a[r] = K[ic];
        b[r] = K[ic ? ic - 1 : 0];
    }
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < RUN_TILE / 256; r++) {
        const uint64_t i = tile + (uint64_t) r * 256 + threadIdx.x;
        c += (i < n && (i == 0 || a[r] != b[r])) ? 1u : 0u;
    }
    for (int d = 32; d > 0; d >>= 1)
        c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0)
        ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        tcnt[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// starts[j] = index of the first key of run j, in key order (toff: exclusive
// prefix of the tile counts); starts[nruns] = n.
template <typename KT>
__global__ void __launch_bounds__(256) k_run_starts(const KT *K, uint64_t n, const uint32_t *toff,
                                                    uint32_t *starts, uint32_t nruns)
{
    __shared__ uint32_t ws[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t tile = (uint64_t) blockIdx.x * RUN_TILE;
    uint32_t base = toff[blockIdx.x];
    KT a[RUN_TILE / 256], b[RUN_TILE / 256];   // as in k_run_counts
#pragma unroll
    for (int r = 0; r < RUN_TILE / 256; r++) {
        const uint64_t i = tile + (uint64_t) r * 256 + threadIdx.x;
        const uint64_t ic = i < n ? i : n - 1;
        a[r] = K[ic];
        b[r] = K[ic ? ic - 1 : 0];
    }
#pragma unroll
    for (int r = 0; r < RUN_TILE / 256; r++) {
        const uint64_t i = tile + (uint64_t) r * 256 + threadIdx.x;
        const bool head = i < n && (i == 0 || a[r] != b[r]);
        const uint64_t m = __ballot(head);
        if (lane == 0)
            ws[wid] = (uint32_t) __popcll(m);
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int w = 0; w < 4; w++) {
            pre += w < wid ? ws[w] : 0u;
            tot += ws[w];
        }
        if (head)
            starts[base + pre + (uint32_t) __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t) i;
        base += tot;
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        starts[nruns] = (uint32_t) n;   // sentinel
}

__global__ void k_zero_u32_sa(uint32_t *p, uint32_t count)
{
    if (threadIdx.x < count)
        p[threadIdx.x] = 0;
}

__global__ void k_compact_heads(const uint32_t *flags, const uint32_t *pos, uint64_t n,
                                uint32_t *starts)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i])
        starts[pos[i]] = (uint32_t) i;
    if (i == 0)
        starts[pos[n - 1] + flags[n - 1]] = (uint32_t) n;   // sentinel
}

// Size class of every bucket run, or CLS_DONE when it needs no sort (one
// element, or complete identical suffixes already in position order).
__global__ void k_bucket_class(const uint32_t *BK, const uint32_t *starts, uint32_t nruns,
                               unsigned K, uint32_t lcap, uint8_t *cls)
{
    uint64_t k = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nruns)
        return;
    uint32_t a = starts[k], size = starts[k + 1] - a;
    bool done = size <= 1 || (BK[a] & 15u) < K;
    cls[k] = done ? CLS_DONE : (uint8_t) seg_class(size, lcap);
}

// Runs per size class per tile of RUN_TILE runs, class-major: tcnt[c * ntiles +
// tile] for the classes c < CLS_BIG + 1 that are sorted or split (CLS_DONE runs
// are left out), so that one exclusive scan of the array gives every (class,
// tile) its place in the class lists laid one after the other.
__global__ void __launch_bounds__(256) k_class_counts(const uint8_t *cls, uint32_t nruns,
                                                      uint32_t ntiles, uint32_t *tcnt)
{
    __shared__ uint32_t ws[4][NCLS];
    const uint64_t tile = (uint64_t) blockIdx.x * RUN_TILE;
    uint32_t c[NCLS] = {};
    for (int r = 0; r < RUN_TILE / 256; r++) {
        const uint64_t k = tile + (uint64_t) r * 256 + threadIdx.x;
        const uint32_t x = k < nruns ? cls[k] : (uint32_t) CLS_DONE;
#pragma unroll
        for (int q = 0; q < NCLS; q++)
            c[q] += x == (uint32_t) q ? 1u : 0u;
    }
#pragma unroll
    for (int q = 0; q < NCLS; q++) {
        for (int d = 32; d > 0; d >>= 1)
            c[q] += __shfl_xor(c[q], d, 64);
        if ((threadIdx.x & 63) == 0)
            ws[threadIdx.x >> 6][q] = c[q];
    }
    __syncthreads();
    if (threadIdx.x < NCLS)
        tcnt[(uint64_t) threadIdx.x * ntiles + blockIdx.x] =
            ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
}

// Every run of a class c < NCLS as (start, size) at its place in the class lists
// (toff: exclusive scan of k_class_counts' array), in run order within a class.
__global__ void __launch_bounds__(256) k_class_scatter(const uint8_t *cls, const uint32_t *starts,
                                                       uint32_t nruns, uint32_t ntiles,
                                                       const uint32_t *toff, uint2 *out)
{
    __shared__ uint32_t ws[4][NCLS];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t tile = (uint64_t) blockIdx.x * RUN_TILE;
    uint32_t base[NCLS];
#pragma unroll
    for (int q = 0; q < NCLS; q++)
        base[q] = toff[(uint64_t) q * ntiles + blockIdx.x];
    for (int r = 0; r < RUN_TILE / 256; r++) {
        const uint64_t k = tile + (uint64_t) r * 256 + threadIdx.x;
        const uint32_t x = k < nruns ? cls[k] : (uint32_t) CLS_DONE;
        uint64_t m[NCLS];
#pragma unroll
        for (int q = 0; q < NCLS; q++) {
            m[q] = __ballot(x == (uint32_t) q);
            if (lane == 0)
                ws[wid][q] = (uint32_t) __popcll(m[q]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NCLS; q++) {
            uint32_t pre = 0, tot = 0;
            for (int w = 0; w < 4; w++) {
                pre += w < wid ? ws[w][q] : 0u;
                tot += ws[w][q];
            }
            if (x == (uint32_t) q)
                out[base[q] + pre + (uint32_t) __popcll(m[q] & ((1ull << lane) - 1ull))] =
                    make_uint2(starts[k], starts[k + 1] - starts[k]);
            base[q] += tot;
        }
        __syncthreads();
    }
}

// ---- oversized runs: every run of one level at once -------------------------------
//
// The elements of all oversized runs of a level are concatenated (run j from
// offset off_j on), keyed by their 29-code chunk at the level's offset, sorted
// stably by that key and then stably by run: every run's elements end up in
// chunk-key order, ties in their previous order, as a radix sort of each run
// alone would leave them.  They go back to the run's range of the suffix array;
// the runs of equal chunk keys still tied past the chunk go to the next level.
// run[j] = (start in the slice, size, off_j, 0).

template <typename R>
__global__ void __launch_bounds__(256) k_big_load(Geom G, const uint64_t *t2, const R *sa,
                                                  uint32_t base, const uint4 *run, uint32_t nrun,
                                                  uint32_t *rid, R *pos, uint64_t *korig,
                                                  uint64_t *K, uint32_t *E)
{
    for (uint32_t j = blockIdx.x; j < nrun; j += gridDim.x) {
        const uint4 r = run[j];
        for (uint32_t i = threadIdx.x; i < r.y; i += blockDim.x) {
            const uint32_t e = r.z + i;
            const R v = sa[r.x + i];   // with its BWT symbol, for the rows no sort writes
            const uint64_t g = sa_pos<R>(v);
            const uint64_t k = chunk_key(t2, g, suf_rem(G, g), base);
            rid[e] = j;
            pos[e] = v;
            korig[e] = k;
            K[e] = k;
            E[e] = e;
        }
    }
}

// Run of every element in chunk-key order: the key of the stable sort by run.
__global__ void k_big_rids(const uint32_t *E, const uint32_t *rid, uint64_t T, uint32_t *RK)
{
    uint64_t o = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (o < T)
        RK[o] = rid[E[o]];
}

// Sorted element o back to its run's range of the suffix array; its chunk key;
// head flag of the runs of equal (run, chunk key).
template <typename R>
__global__ void k_big_store(const uint32_t *E, const uint32_t *rid, const R *pos,
                            const uint64_t *korig, const uint4 *run, uint64_t T, R *sa,
                            uint64_t *KS, uint32_t *flags)
{
    uint64_t o = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= T)
        return;
    const uint32_t e = E[o];
    const uint4 r = run[rid[e]];
    sa[r.x + (o - r.z)] = pos[e];
    const uint64_t k = korig[e];
    KS[o] = k;
    flags[o] = (o == r.z || korig[E[o - 1]] != k) ? 1u : 0u;
}

__global__ void k_big_subruns(const uint32_t *starts, uint32_t nsub, const uint32_t *E,
                              const uint32_t *rid, const uint4 *run, const uint64_t *KS,
                              SegLists out)
{
    uint64_t k = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = k < nsub;
    uint32_t a = valid ? starts[k] : 0;
    uint32_t b = valid ? starts[k + 1] : 0;
    uint32_t at = 0;
    if (valid) {
        const uint4 r = run[rid[E[a]]];
        at = r.x + (a - r.z);
    }
    bool unres = valid && b - a > 1 && (KS[a] & 31ull) == 30ull;
    push_seg(out, unres, at, b - a);
}

// ---- driver ---------------------------------------------------------------------------

// Run starts of equal keys in K[0..n): returns the run count; starts gets
// nruns+1 entries (sentinel n).  flags/pos/starts: n+1 entries each.
template <typename KT>
static int find_runs(pq_ctx_t *ctx, const KT *K, uint64_t n, uint32_t *flags, uint32_t *pos,
                     uint32_t *starts, uint32_t *nruns)
{
    // heads counted per tile, the tile counts scanned, the heads written in order:
    // the keys are read twice, no per-key flag or position array
    const uint64_t ntiles = (n + RUN_TILE - 1) / RUN_TILE;
    if (n == 0) {
        *nruns = 0;
        hipLaunchKernelGGL(k_zero_u32_sa, dim3(1), dim3(64), 0, ctx->stream, starts, 1u);
        PQ_CHECK(hipGetLastError());
        return PQ_OK;
    }
    hipLaunchKernelGGL(k_run_counts<KT>, dim3((unsigned) ntiles), dim3(256), 0, ctx->stream, K, n,
                       flags);
    int rc = pq_scan_u32(ctx, flags, pos, ntiles, nruns);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_run_starts<KT>, dim3((unsigned) ntiles), dim3(256), 0, ctx->stream, K, n,
                       pos, starts, *nruns);
    PQ_CHECK(hipGetLastError());
    return PQ_OK;
}

// Words of 32 codes cached per suffix for runs whose suffixes have at most
// `left` codes after their common prefix (0: too long, sort by chunk levels).
static int words_for(uint32_t left)
{
    return left == 0 ? 1 : left <= 256 ? (int) ((left + 31) / 32) : 0;
}

static uint32_t lcap_for(int W)
{
    return W == 0 ? 0u : lds_cap(W);
}

// Runs per launch: a dispatch covers fewer than 2^32 work-items (64 lanes per
// run for the widest group sort, 256 per run for the LDS sort).
#define SORT_RUNS_PER_LAUNCH (1u << 22)

template <int W, typename R>
static void launch_sorts(pq_ctx_t *ctx, const Geom &G, const uint64_t *t2, R *sa, uint8_t *bwt8,
                         uint32_t base, uint2 *const items[CLS_BIG], const uint32_t cnt[CLS_BIG],
                         const KeyOut &ko)
{
    for (uint32_t at = 0; at < cnt[0]; at += SORT_RUNS_PER_LAUNCH) {
        const uint32_t m = std::min(cnt[0] - at, SORT_RUNS_PER_LAUNCH);
        hipLaunchKernelGGL((k_sort_groups<RUN_SMALL, W, R>),
                           dim3(pq_blocks((uint64_t) m * RUN_SMALL, 256)), dim3(256), 0,
                           ctx->stream, G, t2, sa, bwt8, base, items[0] + at, m, ko);
    }
    for (uint32_t at = 0; at < cnt[1]; at += SORT_RUNS_PER_LAUNCH) {
        const uint32_t m = std::min(cnt[1] - at, SORT_RUNS_PER_LAUNCH);
        hipLaunchKernelGGL((k_sort_groups<RUN_HALF, W, R>),
                           dim3(pq_blocks((uint64_t) m * RUN_HALF, 256)), dim3(256), 0,
                           ctx->stream, G, t2, sa, bwt8, base, items[1] + at, m, ko);
    }
    for (uint32_t at = 0; at < cnt[2]; at += SORT_RUNS_PER_LAUNCH) {
        const uint32_t m = std::min(cnt[2] - at, SORT_RUNS_PER_LAUNCH);
        hipLaunchKernelGGL((k_sort_groups<RUN_WAVE, W, R>),
                           dim3(pq_blocks((uint64_t) m * RUN_WAVE, 256)), dim3(256), 0,
                           ctx->stream, G, t2, sa, bwt8, base, items[2] + at, m, ko);
    }
    for (uint32_t at = 0; at < cnt[3]; at += SORT_RUNS_PER_LAUNCH) {
        const uint32_t m = std::min(cnt[3] - at, SORT_RUNS_PER_LAUNCH);
        hipLaunchKernelGGL((k_sort_wave<2, W, R>), dim3(pq_blocks((uint64_t) m * 64, 256)),
                           dim3(256), 0, ctx->stream, G, t2, sa, bwt8, base, items[3] + at, m,
                           ko);
    }
    for (uint32_t at = 0; at < cnt[4]; at += SORT_RUNS_PER_LAUNCH) {
        const uint32_t m = std::min(cnt[4] - at, SORT_RUNS_PER_LAUNCH);
        hipLaunchKernelGGL((k_sort_wave<4, W, R>), dim3(pq_blocks((uint64_t) m * 64, 256)),
                           dim3(256), 0, ctx->stream, G, t2, sa, bwt8, base, items[4] + at, m,
                           ko);
    }
    for (uint32_t at = 0; at < cnt[5]; at += SORT_RUNS_PER_LAUNCH) {
        const uint32_t m = std::min(cnt[5] - at, SORT_RUNS_PER_LAUNCH);
        if constexpr (lds_cap(W) <= 512)
            // up to 512 suffixes: eight per lane of one wave
            hipLaunchKernelGGL((k_sort_wave<8, W, R>), dim3(pq_blocks((uint64_t) m * 64, 256)),
                               dim3(256), 0, ctx->stream, G, t2, sa, bwt8, base, items[5] + at, m,
                               ko);
        else
            hipLaunchKernelGGL((k_sort_lds<W, R>), dim3(m), dim3(256), 0, ctx->stream, G, t2, sa,
                               bwt8, base, items[5] + at, ko);
    }
}

template <typename R>
static int sort_lists(pq_ctx_t *ctx, const Geom &G, const uint64_t *t2, R *sa, uint8_t *bwt8,
                      uint32_t base, uint2 *const items[CLS_BIG], const uint32_t cnt[CLS_BIG],
                      const KeyOut &ko)
{
    uint32_t left = ctx->Lmax > base ? ctx->Lmax - base : 0;
    switch (words_for(left)) {
    case 1: launch_sorts<1, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 2: launch_sorts<2, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 3: launch_sorts<3, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 4: launch_sorts<4, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 5: launch_sorts<5, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 6: launch_sorts<6, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 7: launch_sorts<7, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    case 8: launch_sorts<8, R>(ctx, G, t2, sa, bwt8, base, items, cnt, ko); break;
    default:
        for (int c = 0; c < CLS_BIG; c++)
            if (cnt[c])
                return ctx->fail(PQ_ERR_STATE, "run classified for cached words it cannot hold");
    }
    PQ_CHECK(hipGetLastError());
    return PQ_OK;
}

// Bucket width for N suffixes: enough codes for about 8 suffixes per bucket,
// then as many more as fit in the same number of 8-bit radix passes over the
// 2K+4-bit key (finer buckets cost nothing there and make the sorts cheaper).
static unsigned bucket_codes(uint64_t N)
{
    unsigned K = 4;
    while (K < 14 && (N >> (2 * K)) > 8)
        K++;
    unsigned passes = (2 * K + 4 + 7) / 8;
    return std::min(14u, (8 * passes - 4) / 2);
}

// Buffers of one slice sort: keys BK (sorted), the slice's range of the suffix
// array and of the BWT bytes, and run scratch (flags and pos alias the radix
// ping-pong buffers, free once the passes are done).
template <typename R> struct SliceSort {
    uint32_t *BK, *flags, *pos, *starts, *counts;
    uint8_t *cls;
    R *sa;
    uint8_t *bwt8;
};

// Sort the ns suffixes of one slice whose bucket keys are sorted in BK and
// whose positions sit in key order in S.sa: runs of equal keys by size class,
// then oversized runs level by level over 29-code chunks.
template <typename R>
static int sort_runs(pq_ctx_t *ctx, const Geom &G, const uint64_t *t2, unsigned K, uint64_t ns,
                     const SliceSort<R> &S, const KeyOut &ko)
{
    int rc = PQ_OK;
    // bucket runs, classified by size with scans
    uint32_t nruns = 0;
    if ((rc = find_runs<uint32_t>(ctx, S.BK, ns, S.flags, S.pos, S.starts, &nruns)))
        return rc;
    uint32_t lcap0 = lcap_for(words_for(ctx->Lmax > K ? ctx->Lmax - K : 0));
    hipLaunchKernelGGL(k_bucket_class, dim3(pq_blocks(nruns, 256)), dim3(256), 0, ctx->stream,
                       S.BK, S.starts, nruns, K, lcap0, S.cls);
    // the class lists one after the other in one buffer (together at most nruns):
    // runs per class and tile, one scan of those counts, one scatter
    uint32_t cnt[NCLS] = {};
    uint2 *lists[NCLS];
    uint2 *lbuf = nullptr;
    if ((rc = pq_slot(ctx, SLOT_LISTS, ((size_t) nruns + NCLS) * sizeof(uint2), &lbuf)))
        return rc;
    {
        const uint32_t ntiles = (uint32_t) (((uint64_t) nruns + RUN_TILE - 1) / RUN_TILE);
        const uint64_t nc = (uint64_t) ntiles * NCLS;
        uint32_t first[NCLS + 1];
        if (ntiles) {
            hipLaunchKernelGGL(k_class_counts, dim3(ntiles), dim3(256), 0, ctx->stream, S.cls,
                               nruns, ntiles, S.flags);
            uint32_t total = 0;
            if ((rc = pq_scan_u32(ctx, S.flags, S.pos, nc, &total)))
                return rc;
            hipLaunchKernelGGL(k_class_scatter, dim3(ntiles), dim3(256), 0, ctx->stream, S.cls,
                               S.starts, nruns, ntiles, S.pos, lbuf);
            PQ_CHECK(hipGetLastError());
            // each class's first place (the scan at its first tile) back on the host
            for (int c = 0; c < NCLS; c++)
                PQ_CHECK(hipMemcpyAsync(&first[c], S.pos + (uint64_t) c * ntiles, 4,
                                        hipMemcpyDeviceToHost, ctx->stream));
            PQ_CHECK(hipStreamSynchronize(ctx->stream));
            first[NCLS] = total;
        } else {
            for (int c = 0; c <= NCLS; c++)
                first[c] = 0;
        }
        for (int c = 0; c < NCLS; c++) {
            lists[c] = lbuf + first[c];
            cnt[c] = first[c + 1] - first[c];
        }
    }
    // sort every bucket past its first K codes
    if ((rc = sort_lists<R>(ctx, G, t2, S.sa, S.bwt8, K, lists, cnt, ko)))
        return rc;
    // deeper levels: their runs share more than the bucket's codes; their keys come
    // from the text
    KeyOut ko_text = ko;
    ko_text.from_text = 1;
    // oversized buckets: MSD over 29-code chunks, level by level
    std::vector<uint2> big(cnt[CLS_BIG]);
    if (cnt[CLS_BIG]) {
        PQ_CHECK(hipMemcpyAsync(big.data(), lists[CLS_BIG], cnt[CLS_BIG] * sizeof(uint2),
                                hipMemcpyDeviceToHost, ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
    }
    uint32_t base = K;
    SegLists SL;
    std::vector<uint4> hrun;
    while (!big.empty() && rc == PQ_OK) {
        const uint32_t nb = (uint32_t) big.size();
        hrun.resize(nb);
        uint64_t T = 0;
        for (uint32_t j = 0; j < nb; j++) {
            hrun[j] = make_uint4(big[j].x, big[j].y, (uint32_t) T, 0u);
            T += big[j].y;
        }
        uint4 *run = nullptr;
        uint64_t *KA = nullptr;
        char *VA = nullptr;
        if ((rc = pq_slot(ctx, SLOT_BIGRUNS, (size_t) nb * sizeof(uint4), &run)) ||
            (rc = pq_slot(ctx, SLOT_BIGK, T * 8 * 4 + 64, &KA)) ||
            (rc = pq_slot(ctx, SLOT_BIGV, T * (5 * 4 + sizeof(R)) + 64, &VA)))
            return rc;
        uint64_t *Kc = KA, *Ka = KA + T, *korig = KA + 2 * T, *KS = KA + 3 * T;
        R *pos = (R *) VA;   // first: 8-byte aligned for 64-bit positions
        uint32_t *rid = (uint32_t *) (pos + T), *Ec = rid + T, *Ea = Ec + T, *RKc = Ea + T,
                 *RKa = RKc + T;
        PQ_CHECK(hipMemcpyAsync(run, hrun.data(), (size_t) nb * sizeof(uint4),
                                hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_big_load<R>, dim3(std::min<uint32_t>(nb, 65536)), dim3(256), 0,
                           ctx->stream, G, t2, (const R *) S.sa, base, run, nb, rid, pos, korig,
                           Kc, Ec);
        // stable by chunk key, then stable by run
        for (unsigned p = 0; p < 8 && rc == PQ_OK; p++) {
            rc = pq_radix_pass(ctx, Kc, Ec, Ka, Ea, T, 8 * p);
            std::swap(Kc, Ka);
            std::swap(Ec, Ea);
        }
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_big_rids, dim3(pq_blocks(T, 256)), dim3(256), 0, ctx->stream, Ec,
                           rid, T, RKc);
        for (unsigned p = 0; p < 4 && (p == 0 || ((nb - 1) >> (8 * p))); p++) {
            if ((rc = pq_radix_pass(ctx, RKc, Ec, RKa, Ea, T, 8 * p)))
                return rc;
            std::swap(RKc, RKa);
            std::swap(Ec, Ea);
        }
        hipLaunchKernelGGL(k_big_store<R>, dim3(pq_blocks(T, 256)), dim3(256), 0, ctx->stream, Ec,
                           rid, (const R *) pos, korig, run, T, S.sa, KS, S.flags);
        uint32_t nsub = 0;
        if ((rc = pq_scan_u32(ctx, S.flags, S.pos, T, &nsub)))
            return rc;
        hipLaunchKernelGGL(k_compact_heads, dim3(pq_blocks(T, 256)), dim3(256), 0, ctx->stream,
                           S.flags, S.pos, T, S.starts);
        // every run pushed from this level holds at least two of its elements
        const uint32_t scap = (uint32_t) (T / 2 + 2);
        uint2 *sbuf = nullptr;
        if ((rc = pq_slot(ctx, SLOT_SUBLISTS, (size_t) scap * NCLS * sizeof(uint2), &sbuf)))
            return rc;
        for (int c = 0; c < NCLS; c++) {
            SL.items[c] = sbuf + (size_t) c * scap;
            SL.cap[c] = scap;
        }
        SL.count = S.counts;
        SL.overflow = S.counts + NCLS;
        const uint32_t nbase = base + CHUNK;
        SL.lcap = lcap_for(words_for(ctx->Lmax > nbase ? ctx->Lmax - nbase : 0));
        PQ_CHECK(hipMemsetAsync(S.counts, 0, 64, ctx->stream));
        hipLaunchKernelGGL(k_big_subruns, dim3(pq_blocks(nsub, 256)), dim3(256), 0, ctx->stream,
                           S.starts, nsub, Ec, rid, run, KS, SL);
        PQ_CHECK(hipGetLastError());
        uint32_t c5[NCLS + 1];
        PQ_CHECK(hipMemcpyAsync(c5, S.counts, (NCLS + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
        if (c5[NCLS])
            return ctx->fail(PQ_ERR_LIMIT, "suffix sort run list overflow");
        base = nbase;
        if ((rc = sort_lists<R>(ctx, G, t2, S.sa, S.bwt8, base, SL.items, c5, ko_text)))
            return rc;
        big.assign(c5[CLS_BIG], make_uint2(0, 0));
        if (c5[CLS_BIG]) {
            PQ_CHECK(hipMemcpyAsync(big.data(), SL.items[CLS_BIG], c5[CLS_BIG] * sizeof(uint2),
                                    hipMemcpyDeviceToHost, ctx->stream));
            PQ_CHECK(hipStreamSynchronize(ctx->stream));
        }
        if (base > ctx->Lmax + CHUNK)
            return ctx->fail(PQ_ERR_STATE, "suffix sort did not converge");
    }
    return rc;
}

// Slices of the suffix set: contiguous key ranges [lo, hi) of at most cap
// suffixes each, cut at SLICE_BINS-histogram boundaries; one slice when N fits.
struct Slice {
    uint64_t lo, hi, count;
};

static int plan_slices(pq_ctx_t *ctx, const Geom &G, const uint64_t *t2, unsigned K, uint64_t cap,
                       std::vector<Slice> &out)
{
    const uint64_t N = ctx->N;
    out.clear();
    if (N <= cap) {
        out.push_back({0, 1ull << 32, N});
        return PQ_OK;
    }
    const unsigned shift = 2 * K + 4 - 12;
    unsigned long long *hist = nullptr;
    int rc;
    if ((rc = pq_slot(ctx, SLOT_SLICE_HIST, SLICE_BINS * 8, &hist)))
        return rc;
    PQ_CHECK(hipMemsetAsync(hist, 0, SLICE_BINS * 8, ctx->stream));
    hipLaunchKernelGGL(k_key_hist, dim3(std::min<unsigned>(1024, pq_blocks((N + 31) / 32, 256))),
                       dim3(256), 0, ctx->stream, G, t2, N, K, shift, hist);
    PQ_CHECK(hipGetLastError());
    std::vector<uint64_t> h(SLICE_BINS);
    PQ_CHECK(hipMemcpyAsync(h.data(), hist, SLICE_BINS * 8, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    uint64_t lo = 0, cur = 0, total = 0;
    for (uint64_t b = 0; b < SLICE_BINS; b++) {
        if (h[b] > cap)
            return ctx->fail(PQ_ERR_LIMIT, "one bucket range holds more suffixes than a slice");
        if (cur + h[b] > cap) {
            out.push_back({lo << shift, b << shift, cur});
            lo = b;
            cur = 0;
        }
        cur += h[b];
        total += h[b];
    }
    if (cur)
        out.push_back({lo << shift, 1ull << 32, cur});
    if (total != N)
        return ctx->fail(PQ_ERR_STATE, "slice histogram does not cover the text");
    return PQ_OK;
}

template <typename R> static int build_sa_dir_t(pq_ctx_t *ctx, int d)
{
    const uint64_t N = ctx->N;
    // the run sorts pack a suffix's position and code count in one word
    if (N >= (1ull << 41) || ctx->Lmax >= (1u << 20))
        return ctx->fail(PQ_ERR_LIMIT, "text of 2^41 codes or reads of 2^20 codes or more");
    DirIndex &D = ctx->dir[d];
    Geom G = ctx->geom();
    const unsigned K = bucket_codes(N);
    const unsigned passes = (2 * K + 4 + 7) / 8;
    uint64_t cap = ctx->opt_slice ? ctx->opt_slice : (1ull << 31);
    cap = std::min<uint64_t>(cap, 1ull << 31);
    std::vector<Slice> slices;
    int rc;
    if ((rc = plan_slices(ctx, G, D.text2, K, cap, slices)))
        return rc;
    ctx->last_slices = (uint32_t) slices.size();
    uint64_t smax = 0;
    for (const Slice &sl : slices)
        smax = std::max(smax, sl.count);
    const uint64_t ntiles = (N + SLICE_TILE - 1) / SLICE_TILE;
    SliceSort<R> S;
    uint32_t *BK2 = nullptr, *bcnt = nullptr;
    R *V2 = nullptr;
    const size_t vbytes = (smax + 1) * std::max<size_t>(4, sizeof(R));
    // several slices (at most SLICE_MAX) with an even number of radix passes: all
    // slices' pairs are made in one pass over the text, the keys into a key array
    // for the whole text (the positions go to the suffix array, where each slice's
    // last radix pass leaves them anyway); else slice by slice
    const bool emit_all = slices.size() > 1 && slices.size() <= SLICE_MAX && !(passes & 1);
    uint32_t *BKall = nullptr;
    // (the whole text's keys live in the search's candidate slot, which the build no
    // longer needs once it is done and the search reuses: the search sizes its
    // batches by free memory, counting that slot as free)
    if ((rc = emit_all ? pq_slot(ctx, SLOT_CAND_LC, N * 4 + 4, &BKall)
                       : pq_slot(ctx, SLOT_BK, smax * 4 + 4, &BKall)) ||
        (rc = pq_slot(ctx, SLOT_BK2, (smax + 1) * 4, &BK2)) ||
        (rc = pq_slot(ctx, SLOT_V2, vbytes, &V2)) ||
        (rc = pq_slot(ctx, SLOT_STARTS, (smax + 2) * 4, &S.starts)) ||
        (rc = pq_slot(ctx, SLOT_CLS, smax + 1, &S.cls)) ||
        (rc = pq_slot(ctx, SLOT_COUNTS, 64, &S.counts)) ||
        (rc = pq_slot(ctx, SLOT_BWT8, N + 1, &D.bwt8)) ||
        (slices.size() > 1 &&
         (rc = pq_slot(ctx, SLOT_SLICE_TILES, (2 * slices.size() + 1) * (ntiles + 1) * 4,
                       &bcnt))))
        return rc;
    // every slice's suffixes per tile, in one pass (up to SLICE_MAX slices)
    const bool count_all = slices.size() > 1 && slices.size() <= SLICE_MAX;
    if (count_all) {
        SliceBounds sb;
        for (size_t i = 0; i <= SLICE_MAX; i++)
            sb.b[i] = i < slices.size() ? (uint32_t) std::min<uint64_t>(slices[i].lo, 0xFFFFFFFFull)
                                        : 0xFFFFFFFFu;
        const uint32_t nsl = (uint32_t) slices.size();
        const dim3 grid((unsigned) ntiles);
        if (nsl <= 4)
            hipLaunchKernelGGL(k_slice_count_all<4>, grid, dim3(256), 0, ctx->stream, G, D.text2,
                               N, K, sb, nsl, ntiles, bcnt);
        else if (nsl <= 8)
            hipLaunchKernelGGL(k_slice_count_all<8>, grid, dim3(256), 0, ctx->stream, G, D.text2,
                               N, K, sb, nsl, ntiles, bcnt);
        else
            hipLaunchKernelGGL(k_slice_count_all<16>, grid, dim3(256), 0, ctx->stream, G,
                               D.text2, N, K, sb, nsl, ntiles, bcnt);
        PQ_CHECK(hipGetLastError());
    }
    S.BK = BKall;
    if (emit_all) {
        // each slice's tile offsets (a scan per slice), then the one emit pass
        SliceBounds sb;
        SliceStarts ss;
        uint64_t at = 0;
        for (size_t i = 0; i <= SLICE_MAX; i++)
            sb.b[i] = i < slices.size() ? (uint32_t) std::min<uint64_t>(slices[i].lo, 0xFFFFFFFFull)
                                        : 0xFFFFFFFFu;
        uint32_t *soff = bcnt + slices.size() * ntiles;   // every slice's tile offsets
        for (size_t i = 0; i < slices.size(); i++) {
            ss.start[i] = at;
            uint32_t got = 0;
            if ((rc = pq_scan_u32(ctx, bcnt + i * ntiles, soff + i * ntiles, ntiles, &got)))
                return rc;
            if (got != slices[i].count)
                return ctx->fail(PQ_ERR_STATE, "slice count disagrees with its histogram");
            at += slices[i].count;
        }
        const uint32_t nsl = (uint32_t) slices.size();
        const dim3 grid((unsigned) ntiles);
        if (nsl <= 4)
            hipLaunchKernelGGL((k_slice_emit_all<R, 4>), grid, dim3(256), 0, ctx->stream, G,
                               D.text2, N, K, sb, ss, nsl, ntiles, soff, BKall, (R *) D.sa);
        else if (nsl <= 8)
            hipLaunchKernelGGL((k_slice_emit_all<R, 8>), grid, dim3(256), 0, ctx->stream, G,
                               D.text2, N, K, sb, ss, nsl, ntiles, soff, BKall, (R *) D.sa);
        else
            hipLaunchKernelGGL((k_slice_emit_all<R, 16>), grid, dim3(256), 0, ctx->stream, G,
                               D.text2, N, K, sb, ss, nsl, ntiles, soff, BKall, (R *) D.sa);
        PQ_CHECK(hipGetLastError());
    }
    S.flags = BK2;
    S.pos = (uint32_t *) V2;
    PQ_CHECK(hipMemsetAsync(D.bwt8, BWT_UNSET, N, ctx->stream));
    // the learned index's sampled keys, written by the run sorts as they place their
    // rows, and by k_bwt_fill for the rows no sort placed
    D.keys_stride = D.keys_D = D.keys_pending = 0;
    KeyOut ko = {nullptr, nullptr, 0, 0, std::min<uint32_t>(63, ctx->t), K, 0};
    const uint64_t nkeys = ctx->key_stride ? (N + ctx->key_stride - 1) / ctx->key_stride : 0;
    if (nkeys) {
        if ((rc = pq_slot(ctx, SLOT_SKEY0 + d, (nkeys + 1) * sizeof(u128), &ko.keys)))
            return rc;
        ko.shift = (uint32_t) __builtin_ctz(ctx->key_stride);
    }
    uint64_t off = 0;
    for (const Slice &sl : slices) {
        const uint64_t ns = sl.count;
        S.sa = (R *) D.sa + off;
        S.bwt8 = D.bwt8 + off;
        if (emit_all)
            S.BK = BKall + off;   // this slice's keys, where the one emit pass put them
        // start in the pair that makes the last radix pass write (BK, slice of sa)
        uint32_t *kc = (passes & 1) ? BK2 : S.BK, *ka = (passes & 1) ? S.BK : BK2;
        R *vc = (passes & 1) ? V2 : S.sa, *va = (passes & 1) ? S.sa : V2;
        // 1. bucket keys by the first K codes (and the length below K)
        unsigned p0 = 0;
        if (slices.size() == 1) {
            // the first pass makes the pairs itself
            if ((rc = radix_pass_src(ctx, BucketSrc<R>{G, D.text2, K}, ka, va, N, 0)))
                return rc;
            std::swap(kc, ka);
            std::swap(vc, va);
            p0 = 1;
        } else if (!emit_all) {
            const size_t si = (size_t) (&sl - slices.data());
            uint32_t *boff = bcnt + slices.size() * ntiles;
            uint32_t *scnt = bcnt + (count_all ? si * ntiles : 0);
            if (!count_all)
                hipLaunchKernelGGL(k_slice_count, dim3((unsigned) ntiles), dim3(256), 0,
                                   ctx->stream, G, D.text2, N, K, sl.lo, sl.hi, scnt);
            uint32_t got = 0;
            if ((rc = pq_scan_u32(ctx, scnt, boff, ntiles, &got)))
                return rc;
            if (got != ns)
                return ctx->fail(PQ_ERR_STATE, "slice count disagrees with its histogram");
            hipLaunchKernelGGL(k_slice_emit<R>, dim3((unsigned) ntiles), dim3(256), 0,
                               ctx->stream, G, D.text2, N, K, sl.lo, sl.hi, boff, kc, vc);
        }
        for (unsigned p = p0; p < passes; p++) {
            if ((rc = pq_radix_pass(ctx, kc, vc, ka, va, ns, 8 * p)))
                return rc;
            std::swap(kc, ka);
            std::swap(vc, va);
        }
        // 2-4. runs of equal keys
        ko.BK = S.BK;
        ko.row0 = off;
        if ((rc = sort_runs<R>(ctx, G, D.text2, K, ns, S, ko)))
            return rc;
        off += ns;
    }
    if (off != N)
        return ctx->fail(PQ_ERR_STATE, "slices do not cover the text");
    D.keys_pending = nkeys ? ctx->key_stride : 0;
    return PQ_OK;
}

int pq_build_sa_dir(pq_ctx_t *ctx, int d)
{
    if (ctx->N == 0)
        return PQ_OK;
    return ctx->wide ? build_sa_dir_t<uint64_t>(ctx, d) : build_sa_dir_t<uint32_t>(ctx, d);
}

// ---- BWT occurrence lines + reads_index ---------------------------------------------------

// BWT symbols of the rows no sort wrote (one-suffix buckets, buckets of
// identical complete suffixes, rows placed by the chunk-level radix passes), and
// those rows' sampled keys when the build writes keys (keys != nullptr).
// One lane per 8 rows: their BWT bytes are read as one word, and only the unset
// ones (a few percent) do any work.
template <typename R>
__global__ void k_bwt_fill(Geom G, const uint64_t *t2, R *sa, uint64_t N, uint8_t *bwt8,
                           u128 *keys, uint32_t shift, uint32_t D)
{
    const uint64_t n8 = (N + 7) / 8;
    PQ_GRID_STRIDE_FOR(q, n8)
    {
        const uint64_t r0 = q * 8;
        // bwt8 holds N + 1 bytes: the last word may be partial
        uint64_t w = 0;
        if (r0 + 8 <= N) {
            w = *(const uint64_t *) (bwt8 + r0);
        } else {
            for (uint64_t r = r0; r < N; r++)
                w |= (uint64_t) bwt8[r] << (8 * (r - r0));
        }
        if (r0 + 8 <= N && (w & 0x8080808080808080ull) == 0)
            continue;   // all eight set (a symbol is at most 4, BWT_UNSET has bit 7)
        for (int j = 0; j < 8; j++) {
            const uint64_t row = r0 + (uint64_t) j;
            if (row >= N || ((w >> (8 * j)) & 0xFFu) != BWT_UNSET)
                continue;
            const R v = sa[row];
            const uint64_t g = sa_pos<R>(v);
            if constexpr (sizeof(R) == 8) {
                // the symbol the bucket pass packed into the value; the row's
                // final value without it
                bwt8[row] = (uint8_t) ((uint64_t) v >> SA_SYM_SHIFT);
                sa[row] = (R) g;
            } else {
                bwt8[row] = (uint8_t) bwt_sym(G, t2, g);
            }
            if (keys && !(row & ((1ull << shift) - 1ull)))
                keys[row >> shift] = suffix_key(G, t2, g, D);
        }
    }
}

// One wave per four occurrence lines of 128 rows (grid-stride): the three bit
// planes of each 64-row half by ballots, each line written by one lane with its
// counts of A, C, G, T (no LDS, no block barrier).  Rows past N carry symbol 7.
__global__ void __launch_bounds__(256) k_occ_planes(const uint8_t *bwt8, uint64_t N,
                                                    uint64_t nlines, OccLine *occ, uint4 *cnts)
{
    // four lines per wave and iteration: the eight bytes of a lane loaded together,
    // without a bounds branch (clamped to byte N, which exists; rows past N read
    // as symbol 7), then lane q writes line q
    constexpr int LPW = 4;
    const int lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t) gridDim.x * blockDim.x) >> 6;
    const uint64_t ngroups = (nlines + LPW - 1) / LPW;
    for (uint64_t grp = ((uint64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6; grp < ngroups;
         grp += nwaves) {
        const uint64_t row0 = grp * (LPW * 128) + (uint64_t) lane;
        unsigned sy[2 * LPW];
#pragma unroll
        for (int k = 0; k < 2 * LPW; k++) {
            const uint64_t r = row0 + 64 * (uint64_t) k;
            sy[k] = bwt8[r < N ? r : N];
        }
        uint64_t pa[2 * LPW], pb[2 * LPW], pc[2 * LPW];
#pragma unroll
        for (int k = 0; k < 2 * LPW; k++) {
            const unsigned v = row0 + 64 * (uint64_t) k < N ? sy[k] : 7u;
            pa[k] = __ballot(v & 1u);
            pb[k] = __ballot(v & 2u);
            pc[k] = __ballot(v & 4u);
        }
#pragma unroll
        for (int q = 0; q < LPW; q++) {
            const uint64_t line = grp * LPW + (uint64_t) q;
            if (lane == q && line < nlines) {
                const uint64_t a0 = pa[2 * q], a1 = pa[2 * q + 1], b0 = pb[2 * q],
                               b1 = pb[2 * q + 1], c0 = pc[2 * q], c1 = pc[2 * q + 1];
                OccLine &ln = occ[line];
                ln.p0[0] = a0;
                ln.p0[1] = a1;
                ln.p1[0] = b0;
                ln.p1[1] = b1;
                ln.p2[0] = c0;
                ln.p2[1] = c1;
                cnts[line] = make_uint4(
                    __popcll(a0 & ~b0 & ~c0) + __popcll(a1 & ~b1 & ~c1),
                    __popcll(~a0 & b0 & ~c0) + __popcll(~a1 & b1 & ~c1),
                    __popcll(a0 & b0 & ~c0) + __popcll(a1 & b1 & ~c1),
                    __popcll(~a0 & ~b0 & c0) + __popcll(~a1 & ~b1 & c1));
            }
        }
    }
}

__global__ void k_occ_counts(OccLine *occ, const uint4 *excl, uint64_t nlines)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nlines) {
        uint4 c = excl[i];
        occ[i].cnt[0] = c.x;
        occ[i].cnt[1] = c.y;
        occ[i].cnt[2] = c.z;
        occ[i].cnt[3] = c.w;
    }
}

// reads_index: the read of every '$' row at its '$' rank.  One lane per 64-row
// half line: the '$' rows are the zeros of its three planes (rows past N carry
// symbol 7), about one in 150 rows for 150-code reads.
template <typename R>
__global__ void k_reads_index(Geom G, const R *sa, const OccLine *occ, uint64_t N, uint32_t *ri)
{
    const uint64_t nh = (N + 63) / 64;
    PQ_GRID_STRIDE_FOR(h, nh)
    {
        const uint64_t line = h >> 1;
        const int half = (int) (h & 1);
        uint64_t dollar = ~(occ[line].p0[half] | occ[line].p1[half] | occ[line].p2[half]);
        const uint64_t row0 = h * 64;
        if (N - row0 < 64)
            dollar &= (1ull << (N - row0)) - 1ull;
        while (dollar) {
            const uint64_t row = row0 + (uint64_t) (__ffsll((long long) dollar) - 1);
            dollar &= dollar - 1ull;
            Ranks rk;
            pq_ranks(occ, row, rk);
            ri[rk.r[0]] = (uint32_t) pq_read_of(G, (uint64_t) sa[row]);
        }
    }
}

int pq_build_occ_dir(pq_ctx_t *ctx, int d)
{
    const uint64_t N = ctx->N;
    DirIndex &D = ctx->dir[d];
    Geom G = ctx->geom();
    D.nlines = N / 128 + 1;
    int rc;
    if (!D.occ && (rc = pq_alloc(ctx, (void **) &D.occ, D.nlines * sizeof(OccLine))))
        return rc;
    if (!D.reads_index && (rc = pq_alloc(ctx, (void **) &D.reads_index, (ctx->n + 1) * 4)))
        return rc;
    uint4 *cnts = (uint4 *) ctx->tmp(SLOT_OCC, D.nlines * sizeof(uint4) * 2);
    if (!cnts)
        return ctx->fail(PQ_ERR_NOMEM, "occ scratch");
    uint4 *excl = cnts + D.nlines;
    if (!D.bwt8)
        return ctx->fail(PQ_ERR_STATE, "BWT symbols missing: suffix array not built");
    u128 *keys = D.keys_pending ? (u128 *) ctx->scratch[SLOT_SKEY0 + d].p : nullptr;
    const uint32_t kshift = D.keys_pending ? (uint32_t) __builtin_ctz(D.keys_pending) : 0u;
    const uint32_t kD = std::min<uint32_t>(63, ctx->t);
    if (ctx->wide)
        hipLaunchKernelGGL(k_bwt_fill<uint64_t>, dim3(pq_blocks_gs((N + 7) / 8, 256)), dim3(256), 0,
                           ctx->stream, G, D.text2, (uint64_t *) D.sa, N, D.bwt8, keys,
                           kshift, kD);
    else
        hipLaunchKernelGGL(k_bwt_fill<uint32_t>, dim3(pq_blocks_gs((N + 7) / 8, 256)), dim3(256), 0,
                           ctx->stream, G, D.text2, (uint32_t *) D.sa, N, D.bwt8, keys,
                           kshift, kD);
    if (D.keys_pending) {
        D.keys_stride = D.keys_pending;
        D.keys_D = kD;
        D.keys_pending = 0;
    }
    hipLaunchKernelGGL(k_occ_planes, dim3(pq_blocks_gs((D.nlines + 3) / 4, 4)), dim3(256), 0,
                       ctx->stream, D.bwt8, N, D.nlines, D.occ, cnts);
    uint32_t tot[4];
    if ((rc = pq_scan_u4(ctx, cnts, excl, D.nlines, tot)))
        return rc;
    hipLaunchKernelGGL(k_occ_counts, dim3(pq_blocks(D.nlines, 256)), dim3(256), 0, ctx->stream,
                       D.occ, excl, D.nlines);
    if (ctx->wide)
        hipLaunchKernelGGL(k_reads_index<uint64_t>, dim3(pq_blocks_gs((N + 63) / 64, 256)), dim3(256), 0,
                           ctx->stream, G, (const uint64_t *) D.sa, D.occ, N, D.reads_index);
    else
        hipLaunchKernelGGL(k_reads_index<uint32_t>, dim3(pq_blocks_gs((N + 63) / 64, 256)), dim3(256), 0,
                           ctx->stream, G, (const uint32_t *) D.sa, D.occ, N, D.reads_index);
    PQ_CHECK(hipGetLastError());
    D.bwt8 = nullptr;   // the byte array is scratch shared by both directions
    // the per-line counts are 32-bit: every symbol must fill fewer than 2^32
    // rows (the '$' rows are the n read starts)
    if ((uint64_t) tot[0] + tot[1] + tot[2] + tot[3] + ctx->n != N)
        return ctx->fail(PQ_ERR_LIMIT, "a symbol fills 2^32 rows or more of the BWT");
    // C (index.c:366-372): code counts of the text = BWT counts + reads ending
    // (fw) / starting (bw) with the code, whose last/first code is not a BWT row.
    const uint64_t *fix = d == 0 ? ctx->sfx : ctx->pfx;
    uint64_t cnt[4];
    for (int c = 0; c < 4; c++) {
        cnt[c] = (uint64_t) tot[c] + fix[c];
        if (cnt[c] >= 0xFFFFFFFFull)   // interval sizes are 32-bit
            return ctx->fail(PQ_ERR_LIMIT, "a symbol occurs 2^32 times or more in the text");
    }
    if (d == 0) {
        ctx->C[0] = 0;
        ctx->C[1] = 0;
        for (int c = 1; c <= 4; c++)
            ctx->C[c + 1] = ctx->C[c] + cnt[c - 1];
    } else {
        for (int c = 1; c <= 4; c++)
            if (ctx->C[c + 1] - ctx->C[c] != cnt[c - 1])
                return ctx->fail(PQ_ERR_STATE, "backward BWT counts disagree with C");
    }
    D.cf[0] = 0;
    for (int c = 1; c <= 4; c++)
        D.cf[c] = ctx->C[c] + fix[c - 1];
    D.built = true;
    return PQ_OK;
}

// ---- upload ---------------------------------------------------------------------------------

__global__ void k_mark_ends(const uint64_t *off, uint64_t n, unsigned long long *ends)
{
    uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && off[r + 1] > off[r]) {
        uint64_t p = off[r + 1] - 1;
        atomicOr(&ends[p >> 6], 1ull << (p & 63));
    }
}

__global__ void k_popc_words(const uint64_t *w, uint64_t nw, uint32_t *out)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw)
        out[i] = (uint32_t) __popcll(w[i]);
}

__global__ void k_pack(Geom G, const uint8_t *codes, uint64_t N, uint64_t nwords, int reverse,
                       uint64_t *t2)
{
    uint64_t w = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords)
        return;
    uint64_t x = 0;
    for (int j = 0; j < 32; j++) {
        uint64_t g = w * 32 + j, c = 0;
        if (g < N) {
            uint64_t src = g;
            if (reverse) {
                uint64_t r = pq_read_of(G, g);
                uint64_t a = pq_read_start(G, r), b = a + pq_read_len(G, r);
                src = a + (b - 1 - g);
            }
            c = (uint64_t) (codes[src] - 1) & 3u;
        }
        x |= c << (62 - 2 * j);
    }
    t2[w] = x;
}

void pq_index_release(pq_ctx_t *ctx)
{
    for (int d = 0; d < 2; d++) {
        DirIndex &D = ctx->dir[d];
        pq_free(D.text2);
        pq_free(D.sa);
        pq_free(D.occ);
        pq_free(D.reads_index);
        D = DirIndex();
    }
    pq_free(ctx->d_off);
    pq_free(ctx->d_ends);
    pq_free(ctx->d_end_rank);
    ctx->d_off = nullptr;
    ctx->d_ends = nullptr;
    ctx->d_end_rank = nullptr;
}

extern "C" int pq_upload_reads(pq_ctx_t *ctx, const uint8_t *codes, const uint64_t *off,
                               uint64_t n, uint32_t min_overlap)
{
    if (!ctx || (!codes && n) || !off)
        return PQ_ERR_ARG;
    auto t0 = std::chrono::steady_clock::now();
    pq_index_release(ctx);
    pq_learned_free(ctx);
    ctx->searched = false;
    ctx->edge_cap_auto = 0;
    ctx->n = n;
    ctx->N = off[n];
    ctx->t = min_overlap;
    if (n >= 0xFFFFFFF0ull)
        return ctx->fail(PQ_ERR_LIMIT, "2^32 reads or more");
    // 64-bit rows and positions from 2^32 codes on (or when asked for)
    ctx->wide = ctx->opt_row_bytes == 8 || (ctx->opt_row_bytes == 0 && ctx->N >= 0xFFFFFFF0ull);
    ctx->h_off.assign(off, off + n + 1);
    uint32_t lmin = UINT32_MAX, lmax = 0;
    memset(ctx->sfx, 0, sizeof(ctx->sfx));
    memset(ctx->pfx, 0, sizeof(ctx->pfx));
    for (uint64_t r = 0; r < n; r++) {
        if (off[r + 1] < off[r])
            return ctx->fail(PQ_ERR_ARG, "offsets must be non-decreasing");
        uint32_t len = (uint32_t) (off[r + 1] - off[r]);
        if (len == 0)
            return ctx->fail(PQ_ERR_ARG, "empty read");
        lmin = std::min(lmin, len);
        lmax = std::max(lmax, len);
        uint8_t a = codes[off[r]], b = codes[off[r + 1] - 1];
        if (a < 1 || a > 4 || b < 1 || b > 4)
            return ctx->fail(PQ_ERR_ARG, "codes must be 1..4");
        ctx->pfx[a - 1]++;
        ctx->sfx[b - 1]++;
    }
    ctx->Lmin = n ? lmin : 0;
    ctx->Lmax = lmax;
    ctx->L = (n && lmin == lmax) ? lmin : 0;
    const uint64_t N = ctx->N;
    // two words of slack past the last code: a 64-code window (two 32-code
    // reads) may start at any text position
    uint64_t nwords64 = N / 64 + 2, nwords32 = N / 32 + 4;
    uint8_t *d_codes = nullptr;
    uint32_t *pc = nullptr;
    int rc;
    if ((rc = pq_alloc(ctx, (void **) &ctx->d_off, (n + 1) * 8)) ||
        (rc = pq_alloc(ctx, (void **) &ctx->d_ends, nwords64 * 8)) ||
        (rc = pq_alloc(ctx, (void **) &ctx->d_end_rank, nwords64 * 4)) ||
        (rc = pq_alloc(ctx, (void **) &d_codes, N + 1)) ||
        (rc = pq_alloc(ctx, (void **) &pc, nwords64 * 4)) ||
        (rc = pq_alloc(ctx, (void **) &ctx->dir[0].text2, nwords32 * 8)) ||
        (rc = pq_alloc(ctx, (void **) &ctx->dir[1].text2, nwords32 * 8)) ||
        (rc = pq_alloc(ctx, (void **) &ctx->dir[0].sa, (N + 1) * (ctx->wide ? 8 : 4))) ||
        (rc = pq_alloc(ctx, (void **) &ctx->dir[1].sa, (N + 1) * (ctx->wide ? 8 : 4)))) {
        pq_free(d_codes);
        pq_free(pc);
        return rc;
    }
    PQ_CHECK(hipMemcpyAsync(ctx->d_off, off, (n + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    if (N)
        PQ_CHECK(hipMemcpyAsync(d_codes, codes, N, hipMemcpyHostToDevice, ctx->stream));
    PQ_CHECK(hipMemsetAsync(ctx->d_ends, 0, nwords64 * 8, ctx->stream));
    hipLaunchKernelGGL(k_mark_ends, dim3(pq_blocks(n, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_off, n, (unsigned long long *) ctx->d_ends);
    hipLaunchKernelGGL(k_popc_words, dim3(pq_blocks(nwords64, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_ends, nwords64, pc);
    rc = pq_scan_u32(ctx, pc, ctx->d_end_rank, nwords64, nullptr);
    if (rc == PQ_OK) {
        Geom G = ctx->geom();
        for (int d = 0; d < 2; d++)
            hipLaunchKernelGGL(k_pack, dim3(pq_blocks(nwords32, 256)), dim3(256), 0, ctx->stream,
                               G, d_codes, N, nwords32, d, ctx->dir[d].text2);
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess)
            rc = ctx->fail(PQ_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    pq_free(d_codes);
    pq_free(pc);
    ctx->timing.upload_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}
