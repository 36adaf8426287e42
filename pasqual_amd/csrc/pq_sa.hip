// pq_sa.hip -- suffix-array construction and FM-index build on the device.
//
// Suffix order (the reference's, sufsort.c:154-383 at one thread): suffixes
// run to the end of their read, end of read sorts first, ties by text
// position.  The sort is an MSD refinement over 29-code chunks:
//
//   key_r(g) = the 29 codes starting at g + 29r (2 bits each, zero padding past
//              the read end) << 5  |  min(codes remaining from g + 29r, 30)
//
// Two suffixes with equal key_r and a length field below 30 are identical
// strings (ordered by position); equal keys with field 30 share the chunk and
// go on to chunk r+1.  Round 0 first buckets all N (key_0, g) pairs by the top
// 16 or 24 key bits with stable LSD radix passes, then every bucket and, in
// later rounds, every unresolved run is sorted on its own by (key_r, g):
// runs of <= 16 / <= 64 elements inside a wave (bitonic network over lanes),
// runs of <= 4096 in LDS by one block, larger runs by global radix passes.
// Each sort emits the unresolved sub-runs of the next round, so a round only
// touches suffixes still tied.
//
// The BWT is never stored as bytes: one pass over the SA writes the 3-bit
// symbol planes of every 128-row OccLine with wave ballots, a scan over lines
// adds the counts, and a last pass writes reads_index from the '$' rows.
#include "pq_ctx.h"

#include <algorithm>
#include <cstring>
#include <chrono>

int pq_radix_pass(pq_ctx_t *ctx, const uint64_t *kin, const uint32_t *vin, uint64_t *kout,
                  uint32_t *vout, uint64_t count, unsigned shift);
int pq_scan_u4(pq_ctx_t *ctx, const void *in, void *out, uint64_t count, void *total_host);

#define CHUNK 29

// ---- keys ------------------------------------------------------------------------

__device__ __forceinline__ uint64_t sa_key(const Geom &G, const uint64_t *t2, uint64_t g,
                                           unsigned r)
{
    uint64_t s = g + (uint64_t) CHUNK * r;
    uint64_t f;
    if (G.L) {
        f = G.L - (s % G.L);
    } else {
        uint64_t w = s >> 6;
        unsigned b = (unsigned) (s & 63);
        uint64_t e = G.ends[w] >> b;
        if (b)
            e |= G.ends[w + 1] << (64 - b);
        e &= (1ull << 30) - 1ull;
        f = e ? (uint64_t) (__ffsll((long long) e)) : 30;  // ffs = index of end + 1
    }
    if (f > 30)
        f = 30;
    uint64_t chars = pq_window(t2, s) >> 6;
    if (f < CHUNK)
        chars &= ~((1ull << (2 * (CHUNK - f))) - 1ull);
    return (chars << 5) | f;
}

__global__ void k_key0(Geom G, const uint64_t *t2, uint64_t N, uint64_t *K, uint32_t *V)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        K[i] = sa_key(G, t2, i, 0);
        V[i] = (uint32_t) i;
    }
}

__global__ void k_keys_range(Geom G, const uint64_t *t2, const uint32_t *sa, uint64_t n,
                             unsigned r, uint64_t *K)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        K[i] = sa_key(G, t2, sa[i], r);
}

// ---- segment lists -------------------------------------------------------------------

struct SegLists {
    uint2 *items[4];
    uint32_t *count;         // 4 counters
    uint32_t cap[4];
    uint32_t *overflow;
};

__device__ __forceinline__ int seg_class(uint32_t size)
{
    return size <= 16 ? 0 : size <= 64 ? 1 : size <= 4096 ? 2 : 3;
}

// Wave-aggregated append; every lane of the wave must call it.
__device__ __forceinline__ void push_seg(const SegLists &S, bool valid, uint32_t start,
                                         uint32_t size)
{
    int lane = threadIdx.x & 63;
    int cls = seg_class(size);
    uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < 4; c++) {
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

__device__ __forceinline__ bool kv_less(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb)
{
    return ka < kb || (ka == kb && va < vb);
}

// One group of G lanes per run (G = 16 or 64); bitonic network over lanes.
template <int G>
__global__ void __launch_bounds__(256) k_sort_groups(Geom Gm, const uint64_t *t2, uint32_t *sa,
                                                     const uint64_t *K, unsigned r, int round0,
                                                     const uint2 *segs, uint32_t nseg,
                                                     SegLists out)
{
    uint64_t tid = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t sid = tid / G;
    int li = (int) (threadIdx.x % G);
    int lane = threadIdx.x & 63;
    bool have = sid < nseg;
    uint2 sg = have ? segs[sid] : make_uint2(0, 0);
    bool e = have && (uint32_t) li < sg.y;
    uint32_t g = e ? sa[sg.x + li] : 0xFFFFFFFFu;
    uint64_t key = e ? (round0 ? K[sg.x + li] : sa_key(Gm, t2, g, r)) : ~0ull;
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint64_t pk = __shfl_xor(key, j, 64);
            uint32_t pg = __shfl_xor(g, j, 64);
            bool up = (li & k) == 0;
            bool lower = (li & j) == 0;
            bool pless = kv_less(pk, pg, key, g);
            bool swap = (lower == up) ? pless : kv_less(key, g, pk, pg);
            if (swap) {
                key = pk;
                g = pg;
            }
        }
    }
    if (e)
        sa[sg.x + li] = g;
    uint64_t prev = __shfl_up(key, 1, 64);
    bool head = e && (li == 0 || prev != key);
    uint64_t hm = __ballot(head);
    int gbase = lane - li;
    uint64_t gm = hm >> gbase;
    if (G < 64)
        gm &= (1ull << G) - 1ull;
    uint64_t above = (li + 1 < 64) ? (gm >> (li + 1)) : 0ull;
    uint32_t next = above ? (uint32_t) (li + 1 + __ffsll((long long) above) - 1) : sg.y;
    bool unres = head && (key & 31ull) == 30ull && next - (uint32_t) li > 1;
    push_seg(out, unres, sg.x + (uint32_t) li, next - (uint32_t) li);
}

// One block per run of 65..4096 elements: bitonic sort in LDS.
__global__ void __launch_bounds__(256) k_sort_block(Geom Gm, const uint64_t *t2, uint32_t *sa,
                                                    const uint64_t *K, unsigned r, int round0,
                                                    const uint2 *segs, SegLists out)
{
    __shared__ uint64_t sk[4096];
    __shared__ uint32_t sv[4096];
    uint2 sg = segs[blockIdx.x];
    uint32_t n = sg.y, P = 1;
    while (P < n)
        P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        if (i < n) {
            uint32_t g = sa[sg.x + i];
            sv[i] = g;
            sk[i] = round0 ? K[sg.x + i] : sa_key(Gm, t2, g, r);
        } else {
            sv[i] = 0xFFFFFFFFu;
            sk[i] = ~0ull;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = threadIdx.x; p < P / 2; p += blockDim.x) {
                uint32_t i = 2 * j * (p / j) + (p % j), ip = i + j;
                bool up = (i & k) == 0;
                uint64_t ka = sk[i], kb = sk[ip];
                uint32_t va = sv[i], vb = sv[ip];
                bool gt = kv_less(kb, vb, ka, va);
                if (gt == up) {
                    sk[i] = kb;
                    sk[ip] = ka;
                    sv[i] = vb;
                    sv[ip] = va;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        sa[sg.x + i] = sv[i];
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < n;
        bool head = valid && (i == 0 || sk[i] != sk[i - 1]);
        uint32_t next = i + 1;
        if (head)
            while (next < n && sk[next] == sk[i])
                next++;
        bool unres = head && (sk[i] & 31ull) == 30ull && next - i > 1;
        push_seg(out, unres, sg.x + i, next - i);
    }
}

// ---- run detection for bucket boundaries and large runs ------------------------------

__global__ void k_head_flags(const uint64_t *K, uint64_t n, unsigned shift, uint32_t *flags)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flags[i] = (i == 0 || (K[i] >> shift) != (K[i - 1] >> shift)) ? 1u : 0u;
}

__global__ void k_compact_heads(const uint32_t *flags, const uint32_t *pos, uint64_t n,
                                uint32_t *starts)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i])
        starts[pos[i]] = (uint32_t) i;
}

// Push runs [starts[k], starts[k+1]) of a range at offset `base`.  all_unresolved:
// bucket runs (every run of size > 1 must be sorted); otherwise a run is
// unresolved when its key has length field 30.
__global__ void k_push_runs(const uint64_t *K, const uint32_t *starts, uint32_t nruns,
                            uint64_t n, uint32_t base, int all_unresolved, SegLists out)
{
    uint64_t k = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = k < nruns;
    uint32_t a = valid ? starts[k] : 0;
    uint32_t b = valid ? (k + 1 < nruns ? starts[k + 1] : (uint32_t) n) : 0;
    bool unres = valid && b - a > 1 && (all_unresolved || (K[a] & 31ull) == 30ull);
    push_seg(out, unres, base + a, b - a);
}

// ---- driver ---------------------------------------------------------------------------

struct SegBufs {
    uint2 *items[2][4] = {{nullptr}};
    uint32_t *counts = nullptr;   // [2][4] + overflow
    uint32_t cap[4] = {0};
};

static SegLists lists_of(SegBufs &B, int which)
{
    SegLists S;
    for (int c = 0; c < 4; c++) {
        S.items[c] = B.items[which][c];
        S.cap[c] = B.cap[c];
    }
    S.count = B.counts + 4 * which;
    S.overflow = B.counts + 8;
    return S;
}

// Find runs of equal (key >> shift) in K[0..n) and push them (offset base).
static int push_runs_of(pq_ctx_t *ctx, const uint64_t *K, uint64_t n, unsigned shift,
                        uint32_t base, int all_unres, SegLists out)
{
    uint32_t *flags = (uint32_t *) ctx->tmp(3, n * sizeof(uint32_t) * 3 + 64);
    if (!flags)
        return ctx->fail(PQ_ERR_NOMEM, "run scratch");
    uint32_t *pos = flags + n, *starts = pos + n;
    unsigned nb = pq_blocks(n, 256);
    hipLaunchKernelGGL(k_head_flags, dim3(nb), dim3(256), 0, ctx->stream, K, n, shift, flags);
    uint32_t nruns = 0;
    int rc = pq_scan_u32(ctx, flags, pos, n, &nruns);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_compact_heads, dim3(nb), dim3(256), 0, ctx->stream, flags, pos, n,
                       starts);
    hipLaunchKernelGGL(k_push_runs, dim3(pq_blocks(nruns, 256)), dim3(256), 0, ctx->stream, K,
                       starts, nruns, n, base, all_unres, out);
    PQ_CHECK(hipGetLastError());
    return PQ_OK;
}

// Sort pairs (K[0..n), V[0..n)) by full key with 8 stable passes (result in place).
static int radix_full(pq_ctx_t *ctx, uint64_t *K, uint32_t *V, uint64_t n, uint64_t *K2,
                      uint32_t *V2)
{
    for (unsigned p = 0; p < 8; p++) {
        const uint64_t *ki = (p & 1) ? K2 : K;
        const uint32_t *vi = (p & 1) ? V2 : V;
        uint64_t *ko = (p & 1) ? K : K2;
        uint32_t *vo = (p & 1) ? V : V2;
        int rc = pq_radix_pass(ctx, ki, vi, ko, vo, n, 8 * p);
        if (rc)
            return rc;
    }
    return PQ_OK;
}

int pq_build_sa_dir(pq_ctx_t *ctx, int d)
{
    const uint64_t N = ctx->N;
    DirIndex &D = ctx->dir[d];
    if (N == 0)
        return PQ_OK;
    Geom G = ctx->geom();
    uint64_t *K = nullptr, *K2 = nullptr;
    uint32_t *V2 = nullptr;
    SegBufs B;
    int rc = PQ_OK;
    const uint64_t capdiv[4] = {2, 17, 65, 4097};
    if ((rc = pq_alloc(ctx, (void **) &K, N * 8)) || (rc = pq_alloc(ctx, (void **) &K2, N * 8)) ||
        (rc = pq_alloc(ctx, (void **) &V2, N * 4)) ||
        (rc = pq_alloc(ctx, (void **) &B.counts, 16 * sizeof(uint32_t))))
        goto done;
    for (int c = 0; c < 4; c++) {
        B.cap[c] = (uint32_t) (N / capdiv[c] + 2);
        for (int w = 0; w < 2; w++)
            if ((rc = pq_alloc(ctx, (void **) &B.items[w][c], (size_t) B.cap[c] * sizeof(uint2))))
                goto done;
    }
    {
        // round 0: keys, bucketing by the top bits
        hipLaunchKernelGGL(k_key0, dim3(pq_blocks(N, 256)), dim3(256), 0, ctx->stream, G,
                           D.text2, N, K, D.sa);
        unsigned topbits = N > (1ull << 26) ? 24 : 16;
        unsigned passes = topbits / 8;
        uint64_t *kc = K, *ka = K2;
        uint32_t *vc = D.sa, *va = V2;
        for (unsigned p = 0; p < passes; p++) {
            unsigned shift = 63 - topbits + 8 * p;
            if ((rc = pq_radix_pass(ctx, kc, vc, ka, va, N, shift)))
                goto done;
            std::swap(kc, ka);
            std::swap(vc, va);
        }
        // the sorted pairs are in (kc, vc); make them (K, D.sa)
        K = kc;
        K2 = ka;
        D.sa = vc;
        V2 = va;
        PQ_CHECK(hipMemsetAsync(B.counts, 0, 16 * sizeof(uint32_t), ctx->stream));
        if ((rc = push_runs_of(ctx, K, N, 63 - topbits, 0, 1, lists_of(B, 0))))
            goto done;

        int cur = 0;
        uint32_t maxround = (ctx->Lmax + CHUNK - 1) / CHUNK + 1;
        for (unsigned r = 0; r <= maxround; r++) {
            uint32_t cnt[9];
            PQ_CHECK(hipMemcpyAsync(cnt, B.counts, 9 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                    ctx->stream));
            PQ_CHECK(hipStreamSynchronize(ctx->stream));
            if (cnt[8]) {
                rc = ctx->fail(PQ_ERR_LIMIT, "suffix sort run list overflow");
                goto done;
            }
            uint32_t *c = cnt + 4 * cur;
            if (c[0] + c[1] + c[2] + c[3] == 0)
                break;
            if (r == maxround) {
                rc = ctx->fail(PQ_ERR_STATE, "suffix sort did not converge");
                goto done;
            }
            int nxt = 1 - cur;
            PQ_CHECK(hipMemsetAsync(B.counts + 4 * nxt, 0, 4 * sizeof(uint32_t), ctx->stream));
            SegLists out = lists_of(B, nxt);
            int round0 = r == 0;
            if (c[0])
                hipLaunchKernelGGL(k_sort_groups<16>, dim3(pq_blocks((uint64_t) c[0] * 16, 256)),
                                   dim3(256), 0, ctx->stream, G, D.text2, D.sa, K, r, round0,
                                   B.items[cur][0], c[0], out);
            if (c[1])
                hipLaunchKernelGGL(k_sort_groups<64>, dim3(pq_blocks((uint64_t) c[1] * 64, 256)),
                                   dim3(256), 0, ctx->stream, G, D.text2, D.sa, K, r, round0,
                                   B.items[cur][1], c[1], out);
            if (c[2])
                hipLaunchKernelGGL(k_sort_block, dim3(c[2]), dim3(256), 0, ctx->stream, G,
                                   D.text2, D.sa, K, r, round0, B.items[cur][2], out);
            PQ_CHECK(hipGetLastError());
            if (c[3]) {
                std::vector<uint2> big(c[3]);
                PQ_CHECK(hipMemcpyAsync(big.data(), B.items[cur][3], c[3] * sizeof(uint2),
                                        hipMemcpyDeviceToHost, ctx->stream));
                PQ_CHECK(hipStreamSynchronize(ctx->stream));
                for (const uint2 &s : big) {
                    uint64_t *keys;
                    if (round0) {
                        keys = K + s.x;
                    } else {
                        keys = K;   // K is free after round 0; use its head as key scratch
                        hipLaunchKernelGGL(k_keys_range, dim3(pq_blocks(s.y, 256)), dim3(256), 0,
                                           ctx->stream, G, D.text2, D.sa + s.x, (uint64_t) s.y, r,
                                           keys);
                    }
                    if ((rc = radix_full(ctx, keys, D.sa + s.x, s.y, K2, V2)))
                        goto done;
                    if ((rc = push_runs_of(ctx, keys, s.y, 0, s.x, 0, out)))
                        goto done;
                }
            }
            cur = nxt;
        }
    }
done:
    pq_free(K);
    pq_free(K2);
    pq_free(V2);
    pq_free(B.counts);
    for (int w = 0; w < 2; w++)
        for (int c = 0; c < 4; c++)
            pq_free(B.items[w][c]);
    return rc;
}

// ---- BWT occurrence lines + reads_index ---------------------------------------------------

__device__ __forceinline__ unsigned bwt_sym(const Geom &G, const uint64_t *t2, uint32_t g)
{
    return pq_is_start(G, g) ? 0u : pq_code(t2, (uint64_t) g - 1) + 1u;
}

// 256 threads = 2 lines; each wave builds one 64-row half of a line.
__global__ void __launch_bounds__(256) k_occ_planes(Geom G, const uint64_t *t2, const uint32_t *sa,
                                                    uint64_t N, uint64_t nlines, OccLine *occ,
                                                    uint4 *cnts)
{
    __shared__ uint32_t part[4][4];
    uint64_t row = (uint64_t) blockIdx.x * 256 + threadIdx.x;
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned sym = row < N ? bwt_sym(G, t2, sa[row]) : 7u;
    uint64_t b0 = __ballot(sym & 1u), b1 = __ballot(sym & 2u), b2 = __ballot(sym & 4u);
    uint64_t line = row >> 7;
    int half = (int) ((row >> 6) & 1);
    if (lane == 0 && line < nlines) {
        occ[line].p0[half] = b0;
        occ[line].p1[half] = b1;
        occ[line].p2[half] = b2;
        part[wid][0] = __popcll(b0 & ~b1 & ~b2);
        part[wid][1] = __popcll(~b0 & b1 & ~b2);
        part[wid][2] = __popcll(b0 & b1 & ~b2);
        part[wid][3] = __popcll(~b0 & ~b1 & b2);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint64_t ln = (uint64_t) blockIdx.x * 2 + threadIdx.x;
        int w0 = 2 * threadIdx.x;
        if (ln < nlines)
            cnts[ln] = make_uint4(part[w0][0] + part[w0 + 1][0], part[w0][1] + part[w0 + 1][1],
                                  part[w0][2] + part[w0 + 1][2], part[w0][3] + part[w0 + 1][3]);
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

__global__ void k_reads_index(Geom G, const uint32_t *sa, const OccLine *occ, uint64_t N,
                              uint32_t *ri)
{
    uint64_t row = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N)
        return;
    uint64_t line = row >> 7;
    int half = (int) ((row >> 6) & 1);
    uint64_t a = occ[line].p0[half], b = occ[line].p1[half], c = occ[line].p2[half];
    unsigned j = (unsigned) (row & 63);
    if ((a >> j) & 1 || (b >> j) & 1 || (c >> j) & 1)
        return;
    Ranks rk;
    pq_ranks(occ, row, rk);
    ri[rk.r[0]] = (uint32_t) pq_read_of(G, sa[row]);
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
    uint4 *cnts = (uint4 *) ctx->tmp(2, D.nlines * sizeof(uint4) * 2);
    if (!cnts)
        return ctx->fail(PQ_ERR_NOMEM, "occ scratch");
    uint4 *excl = cnts + D.nlines;
    hipLaunchKernelGGL(k_occ_planes, dim3(pq_blocks(D.nlines * 128, 256)), dim3(256), 0,
                       ctx->stream, G, D.text2, D.sa, N, D.nlines, D.occ, cnts);
    uint32_t tot[4];
    if ((rc = pq_scan_u4(ctx, cnts, excl, D.nlines, tot)))
        return rc;
    hipLaunchKernelGGL(k_occ_counts, dim3(pq_blocks(D.nlines, 256)), dim3(256), 0, ctx->stream,
                       D.occ, excl, D.nlines);
    hipLaunchKernelGGL(k_reads_index, dim3(pq_blocks(N, 256)), dim3(256), 0, ctx->stream, G,
                       D.sa, D.occ, N, D.reads_index);
    PQ_CHECK(hipGetLastError());
    // C (index.c:366-372): code counts of the text = BWT counts + reads ending
    // (fw) / starting (bw) with the code, whose last/first code is not a BWT row.
    const uint64_t *fix = d == 0 ? ctx->sfx : ctx->pfx;
    uint64_t cnt[4];
    for (int c = 0; c < 4; c++)
        cnt[c] = (uint64_t) tot[c] + fix[c];
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
        D.cf[c] = (uint32_t) (ctx->C[c] + fix[c - 1]);
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
    ctx->searched = false;
    ctx->n = n;
    ctx->N = off[n];
    ctx->t = min_overlap;
    if (ctx->N >= 0xFFFFFFF0ull || n >= 0xFFFFFFF0ull)
        return ctx->fail(PQ_ERR_LIMIT, "text of 2^32 codes or more needs 64-bit positions");
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
    uint64_t nwords64 = N / 64 + 2, nwords32 = N / 32 + 2;
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
        (rc = pq_alloc(ctx, (void **) &ctx->dir[0].sa, N * 4 + 4)) ||
        (rc = pq_alloc(ctx, (void **) &ctx->dir[1].sa, N * 4 + 4))) {
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
