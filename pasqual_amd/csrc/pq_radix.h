// pq_radix.h -- one stable LSD radix pass over (key, value) pairs, for any
// source of the pairs (arrays, or pairs generated on the fly by the suffix-array
// build's first pass).
//
// A block owns a tile of THREADS x ROUNDS pairs; each of its waves owns a
// contiguous part and walks it in coalesced rounds of 64.  Within a round a
// lane finds the lanes holding the same 8-bit digit with 8 ballots, and its rank
// among them with one popcount; a per-wave LDS counter per digit carries the
// rank across rounds, so the ranking is stable without any block-wide barrier
// until the tile is complete.  The tile is then staged in LDS in digit order and
// written out so that lanes write consecutive addresses of each digit's run.
#pragma once
#include "pq_ctx.h"

#define RADIX_THREADS 256

// Items per thread of a radix tile: 16, or 8 when key + value take 12 bytes or
// more (the tile is staged in LDS).
template <typename KT, typename VT> constexpr int radix_rounds()
{
    return sizeof(KT) + sizeof(VT) >= 12 ? 8 : 16;
}

// Threads of a scatter block: 512 for 12-byte pairs (a tile of 4096 pairs, 48 KiB
// of LDS: twice the pairs per digit run of a 2048-pair tile for the same work per
// wave), else 256.
template <typename KT, typename VT> constexpr int radix_threads()
{
    return sizeof(KT) + sizeof(VT) == 12 ? 512 : 256;
}

// Pairs read from two arrays.
template <typename KT, typename VT> struct ArraySrc {
    const KT *k;
    const VT *v;
    __device__ __forceinline__ KT key(uint64_t i) const { return k[i]; }
    __device__ __forceinline__ void load(uint64_t i, KT &key, VT &val) const
    {
        key = k[i];
        val = v[i];
    }
};

#define HIST_TILES 8          // tiles counted per block of the histogram kernel

// Digit counts of HIST_TILES consecutive tiles (the scatter's tiles), each wave
// into its own LDS counters (a quarter of the same-counter atomics per copy: the
// digits of a tile repeat a lot, the low byte of a bucket key being mostly its
// length field).  The counts are laid out digit-major for the scan, so a block
// writes each digit's HIST_TILES counters as one 32-byte store instead of 256
// scattered 4-byte stores per tile.
template <typename KT, int ROUNDS, typename Src>
__global__ void __launch_bounds__(RADIX_THREADS) k_radix_hist(Src src, uint64_t count,
                                                             unsigned shift, uint32_t *hist,
                                                             uint32_t nblocks)
{
    constexpr int RADIX_TILE = RADIX_THREADS * ROUNDS;
    __shared__ uint32_t h[4][256];
    const int wid = threadIdx.x >> 6, d = threadIdx.x;
    const uint32_t t0 = blockIdx.x * HIST_TILES;
    uint32_t cnt[HIST_TILES];
#pragma unroll
    for (int g = 0; g < HIST_TILES; g++) {
        for (int i = threadIdx.x; i < 4 * 256; i += RADIX_THREADS)
            (&h[0][0])[i] = 0;
        __syncthreads();
        const uint64_t base = (uint64_t) (t0 + g) * RADIX_TILE;
        // loads at clamped indices, unconditionally: a load under a bounds branch
        // is waited for before the next one is issued
        KT kv[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            const uint64_t i = base + (uint64_t) r * RADIX_THREADS + threadIdx.x;
            kv[r] = src.key(i < count ? i : count - 1);
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; r++)
            if (base + (uint64_t) r * RADIX_THREADS + threadIdx.x < count)
                atomicAdd(&h[wid][(kv[r] >> shift) & 255u], 1u);
        __syncthreads();
        cnt[g] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
        __syncthreads();
    }
    uint32_t *row = hist + (uint64_t) d * nblocks + t0;
    if (t0 + HIST_TILES <= nblocks && ((nblocks | t0) & 3u) == 0) {
        reinterpret_cast<uint4 *>(row)[0] = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
        reinterpret_cast<uint4 *>(row)[1] = make_uint4(cnt[4], cnt[5], cnt[6], cnt[7]);
    } else {
#pragma unroll
        for (int g = 0; g < HIST_TILES; g++)
            if (t0 + g < nblocks)
                row[g] = cnt[g];
    }
}

template <typename KT, typename VT, int ROUNDS, int THREADS, typename Src>
__global__ void __launch_bounds__(THREADS) k_radix_scatter(Src src, KT *kout, VT *vout,
                                                          uint64_t count, unsigned shift,
                                                          const uint32_t *offs, uint32_t nblocks)
{
    constexpr int NW = THREADS / 64;   // waves of the block, each with its own counters
    constexpr int RADIX_TILE = THREADS * ROUNDS;
    constexpr int RADIX_ROUNDS = ROUNDS;
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t tstart[256];
    __shared__ KT sk[RADIX_TILE];
    __shared__ VT sv[RADIX_TILE];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < NW * 256; i += THREADS)
        (&wcnt[0][0])[i] = 0;
    __syncthreads();

    // Blocks go round-robin to the 8 XCDs; the first 8 * (nblocks / 8) blocks take
    // tiles so that each XCD has a contiguous range of them, and the rest keep
    // their own tile.  A 64-byte line of `offs` holds one digit's offsets of 16
    // consecutive tiles, so it is then read through one XCD's L2.
    const uint32_t q8 = nblocks >> 3;
    const uint32_t tb = blockIdx.x < 8 * q8 ? (blockIdx.x & 7u) * q8 + (blockIdx.x >> 3)
                                            : blockIdx.x;
    const uint64_t tile = (uint64_t) tb * RADIX_TILE;
    const uint64_t wbase = tile + (uint64_t) wid * (RADIX_TILE / NW);
    const uint64_t lt = (1ull << lane) - 1ull;
    KT key[RADIX_ROUNDS];
    VT val[RADIX_ROUNDS];
    uint32_t rnk[RADIX_ROUNDS];
    // every load of the wave's quarter first, at clamped indices and without a
    // bounds branch (a load under one is waited for before the next is issued)
#pragma unroll
    for (int r = 0; r < RADIX_ROUNDS; r++) {
        const uint64_t i = wbase + (uint64_t) r * 64 + lane;
        src.load(i < count ? i : count - 1, key[r], val[r]);
    }
#pragma unroll
    for (int r = 0; r < RADIX_ROUNDS; r++) {
        uint64_t i = wbase + (uint64_t) r * 64 + lane;
        bool ok = i < count;
        unsigned d = (unsigned) (key[r] >> shift) & 255u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t before = (uint32_t) __popcll(peers & lt);
        uint32_t base = wcnt[wid][d];
        rnk[r] = base + before;
        // the highest peer lane updates the running count after all lanes read
        if (ok && (peers >> lane) == 1ull)
            wcnt[wid][d] = base + (uint32_t) __popcll(peers);
    }
    __syncthreads();
    // digit totals of the tile and exclusive prefix over waves and digits (thread d
    // of the first four waves: digit d)
    {
        const int d = threadIdx.x;
        __shared__ uint32_t wsum[4];
        uint32_t c[NW], tot = 0, inc = 0;
        if (d < 256) {
#pragma unroll
            for (int w = 0; w < NW; w++) {
                c[w] = wcnt[w][d];
                tot += c[w];
            }
            // block-wide exclusive scan over the 256 digit totals
            inc = tot;
            for (int s = 1; s < 64; s <<= 1) {
                uint32_t o = __shfl_up(inc, s, 64);
                if (lane >= s)
                    inc += o;
            }
            if (lane == 63)
                wsum[wid] = inc;
        }
        __syncthreads();
        if (d < 256) {
            uint32_t pre = 0;
            for (int w = 0; w < wid; w++)
                pre += wsum[w];
            tstart[d] = pre + inc - tot;
        }
        __syncthreads();
        if (d < 256) {
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                wcnt[w][d] = run;
                run += c[w];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RADIX_ROUNDS; r++) {
        uint64_t i = wbase + (uint64_t) r * 64 + lane;
        if (i < count) {
            unsigned d = (unsigned) (key[r] >> shift) & 255u;
            uint32_t p = tstart[d] + wcnt[wid][d] + rnk[r];
            sk[p] = key[r];
            sv[p] = val[r];
        }
    }
    __syncthreads();
    uint64_t valid = count - tile < RADIX_TILE ? count - tile : RADIX_TILE;
    for (uint32_t p = threadIdx.x; p < valid; p += THREADS) {
        KT k = sk[p];
        unsigned d = (unsigned) (k >> shift) & 255u;
        uint64_t o = (uint64_t) offs[(uint64_t) d * nblocks + tb] + (p - tstart[d]);
        kout[o] = k;
        vout[o] = sv[p];
    }
}

// One stable pass by the 8-bit digit at `shift`.  Scratch slot 1 (histogram).
template <typename Src, typename KT, typename VT>
static int radix_pass_src(pq_ctx_t *ctx, Src src, KT *kout, VT *vout, uint64_t count,
                          unsigned shift)
{
    constexpr int ROUNDS = radix_rounds<KT, VT>();
    constexpr int THREADS = radix_threads<KT, VT>();
    constexpr int RADIX_TILE = THREADS * ROUNDS;
    constexpr int HROUNDS = RADIX_TILE / RADIX_THREADS;   // the histogram's tiles are the same
    if (count == 0)
        return PQ_OK;
    if (count > 0xFFFFFFFFull)
        return ctx->fail(PQ_ERR_LIMIT, "radix pass over more than 2^32 items");
    uint32_t nb = (uint32_t) ((count + RADIX_TILE - 1) / RADIX_TILE);
    uint32_t *hist = (uint32_t *) ctx->tmp(1, (size_t) nb * 256 * sizeof(uint32_t) * 2);
    if (!hist)
        return ctx->fail(PQ_ERR_NOMEM, "radix histogram");
    uint32_t *offs = hist + (size_t) nb * 256;
    hipLaunchKernelGGL((k_radix_hist<KT, HROUNDS, Src>), dim3((nb + HIST_TILES - 1) / HIST_TILES),
                       dim3(RADIX_THREADS), 0, ctx->stream, src, count, shift, hist, nb);
    int rc = pq_scan_u32(ctx, hist, offs, (uint64_t) nb * 256, nullptr);
    if (rc)
        return rc;
    hipLaunchKernelGGL((k_radix_scatter<KT, VT, ROUNDS, THREADS, Src>), dim3(nb), dim3(THREADS), 0,
                       ctx->stream, src, kout, vout, count, shift, offs, nb);
    PQ_CHECK(hipGetLastError());
    return PQ_OK;
}
