// pq_api.hip -- context lifetime, device memory, index build entry point and
// read-back of the index in the reference's host layout.
#include "pq_ctx.h"

#include <chrono>
#include <cstring>

void pq_index_release(pq_ctx_t *ctx);

// Device memory held by the suffix-array build between builds is released when
// an allocation outside the build would otherwise fail (see pq_build_index).
static hipError_t malloc_or_release(pq_ctx_t *ctx, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !ctx->building && ctx->build_scratch_held) {
        (void) hipGetLastError();
        pq_release_build_scratch(ctx);
        e = hipMalloc(p, bytes);
    }
    return e;
}

int pq_alloc(pq_ctx_t *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    hipError_t e = malloc_or_release(ctx, p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        (void) hipGetLastError();
        return ctx->fail(PQ_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                           " bytes: " + hipGetErrorString(e));
    }
    return PQ_OK;
}

void pq_free(void *p)
{
    if (p)
        (void) hipFree(p);
}

void *pq_ctx::tmp(int slot, size_t bytes)
{
    if ((size_t) slot >= scratch.size())
        scratch.resize(slot + 1);
    DevBuf &b = scratch[slot];
    if (b.bytes >= bytes)
        return b.p;
    if (b.p) {
        (void) hipStreamSynchronize(stream);
        if (xstream)
            (void) hipStreamSynchronize(xstream);
        (void) hipFree(b.p);
    }
    b.p = nullptr;
    b.bytes = 0;
    size_t want = bytes + bytes / 4 + 256;
    if (malloc_or_release(this, &b.p, want) != hipSuccess) {
        (void) hipGetLastError();
        b.p = nullptr;
        return nullptr;
    }
    b.bytes = want;
    return b.p;
}

extern "C" int pq_dev_alloc(pq_ctx_t *ctx, uint64_t bytes, void **p)
{
    if (!ctx || !p)
        return PQ_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess)
        return ctx->fail(PQ_ERR_HIP, "hipSetDevice");
    return pq_alloc(ctx, p, bytes);
}

extern "C" void pq_dev_free(void *p)
{
    pq_free(p);
}

extern "C" int pq_dev_upload(pq_ctx_t *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (bytes && (!dst || !src)))
        return PQ_ERR_ARG;
    if (bytes) {
        PQ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return PQ_OK;
}

extern "C" int pq_dev_download(pq_ctx_t *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (bytes && (!dst || !src)))
        return PQ_ERR_ARG;
    if (bytes) {
        PQ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return PQ_OK;
}

extern "C" int pq_ctx_create(pq_ctx_t **out, int device)
{
    if (!out)
        return PQ_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void) hipGetLastError();
        return PQ_ERR_HIP;
    }
    if (device < 0 || device >= ndev)
        return PQ_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess)
        return PQ_ERR_HIP;
    pq_ctx_t *ctx = new pq_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PQ_ERR_HIP;
    }
    *out = ctx;
    return PQ_OK;
}

extern "C" void pq_ctx_destroy(pq_ctx_t *ctx)
{
    if (!ctx)
        return;
    (void) hipSetDevice(ctx->device);
    (void) hipStreamSynchronize(ctx->stream);
    pq_index_release(ctx);
    pq_learned_free(ctx);   // search results live in scratch slots, freed below
    pq_search_graph_free(ctx);
    if (ctx->xstream) {
        (void) hipStreamSynchronize(ctx->xstream);
        (void) hipStreamDestroy(ctx->xstream);
    }
    for (DevBuf &b : ctx->scratch)
        pq_free(b.p);
    (void) hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *pq_last_error(const pq_ctx_t *ctx)
{
    return ctx ? ctx->err.c_str() : "no context";
}

extern "C" uint64_t pq_num_reads(const pq_ctx_t *ctx)
{
    return ctx ? ctx->n : 0;
}

extern "C" uint64_t pq_text_len(const pq_ctx_t *ctx)
{
    return ctx ? ctx->N : 0;
}

extern "C" int pq_sync(pq_ctx_t *ctx)
{
    if (!ctx)
        return PQ_ERR_ARG;
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

extern "C" int pq_set_layout(pq_ctx_t *ctx, uint32_t row_bytes, uint64_t slice_suffixes)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (row_bytes != 0 && row_bytes != 4 && row_bytes != 8)
        return ctx->fail(PQ_ERR_ARG, "row_bytes must be 0, 4 or 8");
    ctx->opt_row_bytes = row_bytes;
    ctx->opt_slice = slice_suffixes;
    return PQ_OK;
}

extern "C" int pq_get_layout(pq_ctx_t *ctx, uint32_t *row_bytes, uint32_t *slices)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (row_bytes)
        *row_bytes = ctx->wide ? 8 : 4;
    if (slices)
        *slices = ctx->last_slices;
    return PQ_OK;
}

// Scratch of the suffix-array build (about 21 bytes per suffix of a slice with
// 64-bit rows).  It stays allocated between builds: freeing and allocating tens
// of GiB again costs more than a second on this device (the freed memory is
// cleared before it can be handed out again), so it is released only when an
// allocation outside the build does not fit otherwise.
void pq_release_build_scratch(pq_ctx_t *ctx)
{
    const int slots[] = {SLOT_RADIX, SLOT_BIGK, SLOT_BIGV, SLOT_BK, SLOT_BK2, SLOT_V2,
                         SLOT_FLAGS, SLOT_POS, SLOT_STARTS, SLOT_CLS, SLOT_LISTS, SLOT_SUBLISTS,
                         SLOT_BWT8, SLOT_SLICE_HIST, SLOT_SLICE_TILES, SLOT_OCC};
    (void) hipStreamSynchronize(ctx->stream);
    for (int s : slots) {
        if ((size_t) s < ctx->scratch.size() && ctx->scratch[s].p) {
            pq_free(ctx->scratch[s].p);
            ctx->scratch[s] = DevBuf();
        }
    }
    ctx->build_scratch_held = false;
}

extern "C" int pq_build_index(pq_ctx_t *ctx)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (!ctx->dir[0].text2)
        return ctx->fail(PQ_ERR_STATE, "no reads uploaded");
    PQ_CHECK(hipSetDevice(ctx->device));
    ctx->searched = false;
    pq_learned_free(ctx);
    hipEvent_t ev[5];
    for (auto &e : ev)
        PQ_CHECK(hipEventCreate(&e));
    int rc = PQ_OK;
    struct Building {
        pq_ctx_t *c;
        ~Building() { c->building = false; }
    } guard{ctx};
    ctx->building = true;
    ctx->build_scratch_held = true;
    for (int d = 0; d < 2 && rc == PQ_OK; d++) {
        PQ_CHECK(hipEventRecord(ev[0], ctx->stream));
        rc = pq_build_sa_dir(ctx, d);
        if (rc)
            break;
        PQ_CHECK(hipEventRecord(ev[1], ctx->stream));
        rc = pq_build_occ_dir(ctx, d);
        if (rc)
            break;
        PQ_CHECK(hipEventRecord(ev[2], ctx->stream));
        PQ_CHECK(hipEventSynchronize(ev[2]));
        float a = 0, b = 0;
        (void) hipEventElapsedTime(&a, ev[0], ev[1]);
        (void) hipEventElapsedTime(&b, ev[1], ev[2]);
        ctx->timing.sa_ms[d] = a;
        ctx->timing.occ_ms[d] = b;
    }
    for (auto &e : ev)
        (void) hipEventDestroy(e);
    return rc;
}

// ---- read-back ------------------------------------------------------------------------------

__global__ void k_bwt_bytes(const OccLine *occ, uint64_t N, uint8_t *out)
{
    PQ_GRID_STRIDE_FOR(row, N)
    {
        const OccLine &ln = occ[row >> 7];
        int h = (int) ((row >> 6) & 1);
        unsigned j = (unsigned) (row & 63);
        unsigned s = (unsigned) ((ln.p0[h] >> j) & 1) | (unsigned) (((ln.p1[h] >> j) & 1) << 1) |
                     (unsigned) (((ln.p2[h] >> j) & 1) << 2);
        out[row] = (uint8_t) s;
    }
}

__global__ void k_occ_ref(const OccLine *occ, uint64_t rows, uint4 fix, uint64_t *bitmap,
                          void *sample, int sbytes)
{
    uint64_t w = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= rows)
        return;
    const OccLine &ln = occ[w >> 1];
    int h = (int) (w & 1);
    uint64_t a = ln.p0[h], b = ln.p1[h], c = ln.p2[h];
    uint64_t eq[5] = {~a & ~b & ~c, a & ~b & ~c, ~a & b & ~c, a & b & ~c, ~a & ~b & c};
    uint64_t base[5];
    uint64_t sum = (uint64_t) ln.cnt[0] + ln.cnt[1] + ln.cnt[2] + ln.cnt[3];
    base[0] = (w >> 1) * 128 - sum;
    for (int k = 1; k < 5; k++)
        base[k] = ln.cnt[k - 1];
    if (h) {
        uint64_t a0 = ln.p0[0], b0 = ln.p1[0], c0 = ln.p2[0];
        uint64_t eq0[5] = {~a0 & ~b0 & ~c0, a0 & ~b0 & ~c0, ~a0 & b0 & ~c0, a0 & b0 & ~c0,
                           ~a0 & ~b0 & c0};
        for (int k = 0; k < 5; k++)
            base[k] += (uint64_t) __popcll(eq0[k]);
    }
    const uint32_t fx[5] = {0u, fix.x, fix.y, fix.z, fix.w};
    for (int k = 0; k < 5; k++) {
        bitmap[(uint64_t) k * rows + w] = eq[k] << 1;   // bit j+1 = row 64w+j (j <= 62)
        uint64_t v = base[k] + fx[k];
        if (sbytes == 4)
            ((uint32_t *) sample)[5 * w + k] = (uint32_t) v;
        else
            ((uint64_t *) sample)[5 * w + k] = v;
    }
}

__global__ void k_shortcut(Geom G, const uint64_t *fwt, uint64_t n, int dir, CTab<uint64_t> C,
                           uint64_t *start, uint64_t *size, uint16_t *depth)
{
    uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    uint64_t s = pq_read_start(G, r), L = pq_read_len(G, r);
    unsigned c = pq_code(fwt, dir == 0 ? s + L - 1 : s) + 1;
    start[r] = pq_pick6(C.c, c);
    size[r] = pq_pick6(C.c, c + 1) - pq_pick6(C.c, c);
    depth[r] = 1;
}

// ---- exact interval shortcuts (fixed read length L, N < 2^32) ----------------------------
//
// The reference's MSD sort records, for a suffix of length exactly P (the
// read's last P codes of the direction text) with P = 9, 12, ... <= min(t, L),
// the SA range of the bucket it lies in at depth P, when that bucket is one the
// sort visits: a depth-9 bucket of more than one suffix goes to process_slot_1
// (sufsort.c:594-611), and inside it a child bucket at depth Q + 3 of size f > 1
// recurses into process_slot_1 when f * (L - Q) >= 3500 and is finished by
// insert_sort otherwise (sufsort.c:354-376); both record suffixes of length
// exactly their depth (sufsort.c:177-209, 322-334), the deepest record wins.
// The bucket of row x at depth Q is the run of rows around x whose common
// prefix with their predecessor is at least Q codes.

// Common prefix (in codes, at most cap) of the suffixes of rows i-1 and i.
template <typename R>
__global__ void k_sc_lcp(const R *sa, const uint64_t *t2, uint64_t N, uint32_t L, uint32_t cap,
                         uint16_t *lcp)
{
    PQ_GRID_STRIDE_FOR(i, N)
    {
        if (i == 0) {
            lcp[0] = 0;
            continue;
        }
        const uint64_t a = (uint64_t) sa[i - 1], b = (uint64_t) sa[i];
        const uint32_t ra = L - (uint32_t) (a % L), rb = L - (uint32_t) (b % L);
        uint32_t m = ra < rb ? ra : rb;
        m = m < cap ? m : cap;
        uint32_t k = 0;
        while (k < m) {
            const uint64_t x = pq_window(t2, a + k) ^ pq_window(t2, b + k);
            const uint32_t same = x ? (uint32_t) __clzll(x) / 2 : 32u;
            k += same;
            if (same < 32)
                break;
        }
        lcp[i] = (uint16_t) (k < m ? k : m);
    }
}

// Row of every (read, P) suffix: rowP[r * nP + (P - 9) / 3].
template <typename R>
__global__ void k_sc_rows(const R *sa, uint64_t N, uint32_t L, uint32_t Pmax, uint32_t nP,
                          uint32_t *rowP)
{
    PQ_GRID_STRIDE_FOR(i, N)
    {
        const uint64_t pos = (uint64_t) sa[i];
        const uint32_t P = L - (uint32_t) (pos % L);
        if (P >= 9 && P <= Pmax && (P - 9) % 3 == 0)
            rowP[(pos / L) * nP + (P - 9) / 3] = (uint32_t) i;
    }
}

// Run boundaries at depth Q.
__global__ void k_sc_flags(const uint16_t *lcp, uint64_t N, uint32_t Q, uint32_t *flag)
{
    PQ_GRID_STRIDE_FOR(i, N)
        flag[i] = (i == 0 || lcp[i] < Q) ? 1u : 0u;
}

// start[k] = first row of run k (runs numbered by the exclusive scan of the
// flags); start[runs] = N.
__global__ void k_sc_starts(const uint32_t *flag, const uint32_t *excl, uint64_t N, uint32_t runs,
                            uint32_t *start)
{
    PQ_GRID_STRIDE_FOR(i, N)
    {
        if (flag[i])
            start[excl[i]] = (uint32_t) i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        start[runs] = (uint32_t) N;
}

// Depth Q for every (read, P >= Q): the visit state of its bucket chain
// (vis[r * nP + p]) and, at Q == P, the record.
__global__ void k_sc_level(const uint32_t *rowP, const uint32_t *flag, const uint32_t *excl,
                           const uint32_t *start, uint64_t n, uint32_t nP, uint32_t Q, uint32_t L,
                           uint8_t *vis, uint64_t *istart, uint64_t *isize, uint16_t *idepth)
{
    PQ_GRID_STRIDE_FOR(q, n * nP)
    {
        const uint32_t p = (uint32_t) (q % nP), P = 9 + 3 * p;
        if (P < Q)
            continue;
        const uint32_t x = rowP[q];
        const uint32_t run = excl[x] + flag[x] - 1u;
        const uint32_t rs = start[run], size = start[run + 1] - rs;
        bool v;
        if (Q == 9) {
            v = size > 1;
            vis[q] = v ? 1 : 0;
        } else {
            v = vis[q] && (Q - 3) < L && size > 1;
            vis[q] = (v && (uint64_t) size * (L - (Q - 3)) >= 3500u) ? 1 : 0;
        }
        if (Q == P && v) {            // one writer per read per level; deeper levels later
            const uint64_t r = q / nP;
            istart[r] = rs;
            isize[r] = size;
            idepth[r] = (uint16_t) P;
        }
    }
}

template <typename R>
static int exact_shortcut(pq_ctx_t *ctx, int dir, uint64_t *ds, uint64_t *dz, uint16_t *dd)
{
    const uint64_t n = ctx->n, N = ctx->N;
    const uint32_t L = ctx->L, t = ctx->t;
    const uint32_t Pmax = t < L ? t : L;
    if (Pmax < 9)
        return PQ_OK;
    const uint32_t nP = (Pmax - 9) / 3 + 1;
    uint16_t *lcp = nullptr;
    uint32_t *rowP = nullptr, *flag = nullptr, *excl = nullptr, *start = nullptr;
    uint8_t *vis = nullptr;
    int rc;
    if ((rc = pq_alloc(ctx, (void **) &lcp, N * 2)) ||
        (rc = pq_alloc(ctx, (void **) &rowP, n * nP * 4)) ||
        (rc = pq_alloc(ctx, (void **) &flag, N * 4)) ||
        (rc = pq_alloc(ctx, (void **) &excl, N * 4)) ||
        (rc = pq_alloc(ctx, (void **) &start, (N + 1) * 4)) ||
        (rc = pq_alloc(ctx, (void **) &vis, n * nP))) {
        pq_free(lcp), pq_free(rowP), pq_free(flag), pq_free(excl), pq_free(start), pq_free(vis);
        return rc;
    }
    auto run = [&]() -> int {
        const R *sa = (const R *) ctx->dir[dir].sa;
        const uint64_t *t2 = ctx->dir[dir].text2;
        hipLaunchKernelGGL(k_sc_lcp<R>, dim3(pq_blocks_gs(N, 256)), dim3(256), 0, ctx->stream, sa,
                           t2, N, L, Pmax, lcp);
        hipLaunchKernelGGL(k_sc_rows<R>, dim3(pq_blocks_gs(N, 256)), dim3(256), 0, ctx->stream,
                           sa, N, L, Pmax, nP, rowP);
        PQ_CHECK(hipGetLastError());
        for (uint32_t Q = 9; Q <= Pmax; Q += 3) {
            hipLaunchKernelGGL(k_sc_flags, dim3(pq_blocks_gs(N, 256)), dim3(256), 0, ctx->stream,
                               lcp, N, Q, flag);
            uint32_t runs = 0;
            int rc2 = pq_scan_u32(ctx, flag, excl, N, &runs);
            if (rc2)
                return rc2;
            hipLaunchKernelGGL(k_sc_starts, dim3(pq_blocks_gs(N, 256)), dim3(256), 0,
                               ctx->stream, flag, excl, N, runs, start);
            hipLaunchKernelGGL(k_sc_level, dim3(pq_blocks_gs(n * nP, 256)), dim3(256), 0,
                               ctx->stream, rowP, flag, excl, start, n, nP, Q, L, vis, ds, dz,
                               dd);
            PQ_CHECK(hipGetLastError());
        }
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
        return PQ_OK;
    };
    rc = run();
    pq_free(lcp), pq_free(rowP), pq_free(flag), pq_free(excl), pq_free(start), pq_free(vis);
    return rc;
}

// ---- first-triplet slot sizes (initial_sort, sufsort.c:386-450) ------------------------
//
// The reference codes position g by the (up to) three codes from g to the end of
// its read, 1 + 21 c0 + [rem >= 2] (1 + 5 c1 + [rem >= 3] (1 + c2)) (sufsort.c:
// 21-40, 124-151), and sizes its five sort slots by the codes 17k .. 17k+16.
__global__ void k_slot_hist(Geom G, const uint64_t *t2, unsigned long long *cnt)
{
    __shared__ unsigned long long h[5];
    if (threadIdx.x < 5)
        h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t k0 = 0, k1 = 0, k2 = 0, k3 = 0, k4 = 0;
    PQ_GRID_STRIDE_FOR(g, G.N)
    {
        const uint64_t rem = pq_rem(G, g);
        const uint64_t w = pq_window(t2, g);
        const uint32_t c0 = (uint32_t) (w >> 62), c1 = (uint32_t) (w >> 60) & 3u,
                       c2 = (uint32_t) (w >> 58) & 3u;
        const uint32_t code =
            1u + 21u * c0 + (rem >= 2 ? 1u + 5u * c1 + (rem >= 3 ? 1u + c2 : 0u) : 0u);
        const uint32_t k = code / 17u;
        k0 += k == 0;
        k1 += k == 1;
        k2 += k == 2;
        k3 += k == 3;
        k4 += k == 4;
    }
    atomicAdd(&h[0], (unsigned long long) k0);
    atomicAdd(&h[1], (unsigned long long) k1);
    atomicAdd(&h[2], (unsigned long long) k2);
    atomicAdd(&h[3], (unsigned long long) k3);
    atomicAdd(&h[4], (unsigned long long) k4);
    __syncthreads();
    if (threadIdx.x < 5 && h[threadIdx.x])
        atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_zero_slots(unsigned long long *cnt)
{
    if (threadIdx.x < 5)
        cnt[threadIdx.x] = 0;
}

static int dir_ok(pq_ctx_t *ctx, int dir);

extern "C" int pq_get_slot_sizes(pq_ctx_t *ctx, int dir, uint64_t size[5], uint64_t offset[5],
                                 uint64_t *max_size)
{
    int rc = dir_ok(ctx, dir);
    if (rc)
        return rc;
    if (!size || !offset || !max_size)
        return ctx->fail(PQ_ERR_ARG, "pq_get_slot_sizes: null output");
    unsigned long long *cnt = (unsigned long long *) ctx->tmp(SLOT_MISC2, 64);
    if (!cnt)
        return ctx->fail(PQ_ERR_NOMEM, "slot counters");
    hipLaunchKernelGGL(k_zero_slots, dim3(1), dim3(64), 0, ctx->stream, cnt);
    if (ctx->N)
        hipLaunchKernelGGL(k_slot_hist, dim3(pq_blocks_gs(ctx->N, 256)), dim3(256), 0,
                           ctx->stream, ctx->geom(), ctx->dir[dir].text2, cnt);
    PQ_CHECK(hipGetLastError());
    unsigned long long h[5];
    PQ_CHECK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    *max_size = 0;
    for (int k = 0; k < 5; k++) {
        size[k] = h[k];
        offset[k] = k ? offset[k - 1] + size[k - 1] : 0;
        *max_size = size[k] > *max_size ? size[k] : *max_size;
    }
    return PQ_OK;
}

static int dir_ok(pq_ctx_t *ctx, int dir)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (dir != 0 && dir != 1)
        return ctx->fail(PQ_ERR_ARG, "dir must be 0 or 1");
    if (!ctx->dir[dir].built)
        return ctx->fail(PQ_ERR_STATE, "index not built");
    return PQ_OK;
}

extern "C" int pq_get_sa(pq_ctx_t *ctx, int dir, uint64_t *sa)
{
    int rc = dir_ok(ctx, dir);
    if (rc)
        return rc;
    if (ctx->wide) {
        if (ctx->N)
            PQ_CHECK(hipMemcpy(sa, ctx->dir[dir].sa, ctx->N * 8, hipMemcpyDeviceToHost));
        return PQ_OK;
    }
    std::vector<uint32_t> h(ctx->N);
    if (ctx->N)
        PQ_CHECK(hipMemcpy(h.data(), ctx->dir[dir].sa, ctx->N * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < ctx->N; i++)
        sa[i] = h[i];
    return PQ_OK;
}

extern "C" int pq_get_bwt(pq_ctx_t *ctx, int dir, uint8_t *bwt)
{
    int rc = dir_ok(ctx, dir);
    if (rc)
        return rc;
    if (!ctx->N)
        return PQ_OK;
    uint8_t *d = (uint8_t *) ctx->tmp(4, ctx->N);
    if (!d)
        return ctx->fail(PQ_ERR_NOMEM, "bwt scratch");
    hipLaunchKernelGGL(k_bwt_bytes, dim3(pq_blocks_gs(ctx->N, 256)), dim3(256), 0, ctx->stream,
                       ctx->dir[dir].occ, ctx->N, d);
    PQ_CHECK(hipMemcpyAsync(bwt, d, ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

extern "C" int pq_get_reads_index(pq_ctx_t *ctx, int dir, uint64_t *ri)
{
    int rc = dir_ok(ctx, dir);
    if (rc)
        return rc;
    std::vector<uint32_t> h(ctx->n);
    if (ctx->n)
        PQ_CHECK(hipMemcpy(h.data(), ctx->dir[dir].reads_index, ctx->n * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < ctx->n; i++)
        ri[i] = h[i];
    return PQ_OK;
}

extern "C" int pq_get_counts(pq_ctx_t *ctx, uint64_t C[6], uint64_t suffix_count[4],
                             uint64_t prefix_count[4])
{
    int rc = dir_ok(ctx, 0);
    if (rc)
        return rc;
    if (C)
        memcpy(C, ctx->C, sizeof(ctx->C));
    if (suffix_count)
        memcpy(suffix_count, ctx->sfx, sizeof(ctx->sfx));
    if (prefix_count)
        memcpy(prefix_count, ctx->pfx, sizeof(ctx->pfx));
    return PQ_OK;
}

extern "C" int pq_get_occ_ref(pq_ctx_t *ctx, int dir, uint64_t *bitmap, void *sample,
                              int sample_bytes)
{
    int rc = dir_ok(ctx, dir);
    if (rc)
        return rc;
    if (sample_bytes != 4 && sample_bytes != 8)
        return ctx->fail(PQ_ERR_ARG, "sample_bytes must be 4 or 8");
    uint64_t rows = (ctx->N + 64) / 64;
    size_t bbytes = rows * 5 * 8, sbytes = rows * 5 * (size_t) sample_bytes;
    char *d = (char *) ctx->tmp(4, bbytes + sbytes);
    if (!d)
        return ctx->fail(PQ_ERR_NOMEM, "occ export scratch");
    const uint64_t *fix = dir == 0 ? ctx->sfx : ctx->pfx;
    uint4 f = make_uint4((uint32_t) fix[0], (uint32_t) fix[1], (uint32_t) fix[2], (uint32_t) fix[3]);
    hipLaunchKernelGGL(k_occ_ref, dim3(pq_blocks(rows, 256)), dim3(256), 0, ctx->stream,
                       ctx->dir[dir].occ, rows, f, (uint64_t *) d, (void *) (d + bbytes),
                       sample_bytes);
    PQ_CHECK(hipGetLastError());
    if (bitmap)
        PQ_CHECK(hipMemcpyAsync(bitmap, d, bbytes, hipMemcpyDeviceToHost, ctx->stream));
    if (sample)
        PQ_CHECK(hipMemcpyAsync(sample, d + bbytes, sbytes, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

extern "C" int pq_get_shortcut(pq_ctx_t *ctx, int dir, uint64_t *start, uint64_t *size,
                               uint16_t *depth)
{
    int rc = dir_ok(ctx, dir);
    if (rc)
        return rc;
    uint64_t n = ctx->n;
    if (!n)
        return PQ_OK;
    char *d = (char *) ctx->tmp(4, n * 18 + 64);
    if (!d)
        return ctx->fail(PQ_ERR_NOMEM, "shortcut scratch");
    uint64_t *ds = (uint64_t *) d, *dz = ds + n;
    uint16_t *dd = (uint16_t *) (dz + n);
    hipLaunchKernelGGL(k_shortcut, dim3(pq_blocks(n, 256)), dim3(256), 0, ctx->stream,
                       ctx->geom(), ctx->dir[0].text2, n, dir, ctx->ctab<uint64_t>(), ds, dz, dd);
    PQ_CHECK(hipGetLastError());
    // deeper records where the reference's sort makes them: reads of one length
    // (the reference's only input) and texts of fewer than 2^32 codes
    if (ctx->L && ctx->N < 0xFFFFFFFFull) {
        rc = ctx->wide ? exact_shortcut<uint64_t>(ctx, dir, ds, dz, dd)
                       : exact_shortcut<uint32_t>(ctx, dir, ds, dz, dd);
        if (rc)
            return rc;
    }
    if (start)
        PQ_CHECK(hipMemcpyAsync(start, ds, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (size)
        PQ_CHECK(hipMemcpyAsync(size, dz, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (depth)
        PQ_CHECK(hipMemcpyAsync(depth, dd, n * 2, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

// A plain streaming copy of `bytes` bytes between two device buffers: the
// reference transfer for scaling hardware traffic counters.
__global__ void k_copy_probe(const uint4 *src, uint4 *dst, uint64_t n16)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
    for (; i < n16; i += stride)
        dst[i] = src[i];
}

extern "C" int pq_probe_copy(pq_ctx_t *ctx, uint64_t bytes, double *ms)
{
    if (!ctx || bytes < 16)
        return PQ_ERR_ARG;
    uint64_t n16 = bytes / 16;
    char *buf = (char *) ctx->tmp(SLOT_PROBE, n16 * 32);
    if (!buf)
        return ctx->fail(PQ_ERR_NOMEM, "copy probe buffers");
    hipEvent_t a, b;
    PQ_CHECK(hipEventCreate(&a));
    PQ_CHECK(hipEventCreate(&b));
    PQ_CHECK(hipEventRecord(a, ctx->stream));
    // 1024 blocks of one 16-byte load in flight per lane: the fastest of grids
    // 512-16384 with 1-8 loads in flight per lane on MI355X (1 GiB copies)
    hipLaunchKernelGGL(k_copy_probe, dim3(1024), dim3(256), 0, ctx->stream, (const uint4 *) buf,
                       (uint4 *) (buf + n16 * 16), n16);
    PQ_CHECK(hipEventRecord(b, ctx->stream));
    PQ_CHECK(hipEventSynchronize(b));
    float t = 0;
    (void) hipEventElapsedTime(&t, a, b);
    if (ms)
        *ms = t;
    (void) hipEventDestroy(a);
    (void) hipEventDestroy(b);
    return PQ_OK;
}

// Streaming read of n16 16-byte words.  The xor of what a lane read is
// stored to the sink only when it equals one value, so the loads stay and the
// kernel writes at most 4 bytes whatever the buffer holds.
__global__ void k_read_probe(const uint4 *src, uint64_t n16, uint32_t *sink)
{
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t) gridDim.x * blockDim.x) {
        const uint4 v = src[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9E3779B9u)
        sink[0] = acc;
}

extern "C" int pq_probe_read(pq_ctx_t *ctx, uint64_t bytes, double *ms)
{
    if (!ctx || bytes < 16)
        return PQ_ERR_ARG;
    const uint64_t n16 = bytes / 16;
    char *buf = (char *) ctx->tmp(SLOT_PROBE, n16 * 16 + 64);
    if (!buf)
        return ctx->fail(PQ_ERR_NOMEM, "read probe buffer");
    // 4096 blocks: the fastest grid of streaming reads on MI355X (1 GiB).  One
    // untimed pass first: lines a previous kernel left dirty in the caches would
    // otherwise be written back during the timed one.
    uint32_t *sink = (uint32_t *) (buf + n16 * 16);
    hipLaunchKernelGGL(k_read_probe, dim3(4096), dim3(256), 0, ctx->stream, (const uint4 *) buf,
                       n16, sink);
    hipEvent_t a, b;
    PQ_CHECK(hipEventCreate(&a));
    PQ_CHECK(hipEventCreate(&b));
    PQ_CHECK(hipEventRecord(a, ctx->stream));
    hipLaunchKernelGGL(k_read_probe, dim3(4096), dim3(256), 0, ctx->stream, (const uint4 *) buf,
                       n16, sink);
    PQ_CHECK(hipEventRecord(b, ctx->stream));
    PQ_CHECK(hipEventSynchronize(b));
    float t = 0;
    (void) hipEventElapsedTime(&t, a, b);
    if (ms)
        *ms = t;
    (void) hipEventDestroy(a);
    (void) hipEventDestroy(b);
    return PQ_OK;
}

// Scattered 64-byte line reads over a buffer of nlines OccLines, 2^25 lines
// per launch at hashed positions (below nlines by construction: multiply-high).
// Modes: 0 = every lane loads whole lines as a search step does (four 16-byte
// loads), positions independent of each other (the throughput ceiling of
// scattered lines); 1 = the same with each position hashed from the previous
// line's contents (chains of dependent loads, as in a backward search); 2 =
// the four lanes of a quad load one line together, 16 bytes each; 3 = every
// lane loads the 40 bytes of a line that one rank needs (three loads: 16, 16
// and 8 bytes from byte 0 or 24).  Same sink rule as k_read_probe.
__device__ __forceinline__ uint64_t probe_hash(uint64_t h)
{
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
    return h ^ (h >> 31);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_gather_probe(const OccLine *src, uint64_t nlines,
                                                      uint32_t k, uint32_t *sink)
{
    const uint64_t lane = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t who = MODE == 2 ? lane >> 2 : lane;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < k; i++) {
        const uint64_t h = probe_hash(who * 0x9E3779B97F4A7C15ull + i + (MODE == 1 ? acc : 0u));
        const uint64_t idx = __umul64hi(h, nlines);
        if (MODE == 2) {
            const uint4 v = reinterpret_cast<const uint4 *>(src + idx)[lane & 3];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        } else if (MODE == 3) {
            const uint32_t *w = reinterpret_cast<const uint32_t *>(src + idx) + ((h & 1) ? 6 : 0);
            const uint4 a = *reinterpret_cast<const uint4 *>(w);
            const uint4 b = *reinterpret_cast<const uint4 *>(w + 4);
            const uint2 c = *reinterpret_cast<const uint2 *>(w + 8);
            acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y;
        } else {
            OccLine ln;
            pq_load_line(src, idx, ln);
            acc ^= (uint32_t) (ln.p0[0] ^ ln.p0[1] ^ ln.p1[0] ^ ln.p1[1] ^ ln.p2[0] ^ ln.p2[1]) ^
                   (uint32_t) ((ln.p0[0] ^ ln.p1[1] ^ ln.p2[0]) >> 32) ^ ln.cnt[0] ^ ln.cnt[1] ^
                   ln.cnt[2] ^ ln.cnt[3];
        }
    }
    if (acc == 0x9E3779B9u)
        sink[0] = acc;
}

extern "C" int pq_probe_gather(pq_ctx_t *ctx, uint64_t bytes, int mode, double *ms)
{
    if (!ctx || bytes < 64 || mode < 0 || mode > 3)
        return PQ_ERR_ARG;
    const uint64_t nlines = bytes / 64;
    char *buf = (char *) ctx->tmp(SLOT_PROBE, nlines * 64 + 64);
    if (!buf)
        return ctx->fail(PQ_ERR_NOMEM, "gather probe buffer");
    // 2^20 lanes (4096 blocks: several rounds of the GPU's resident waves) x 32
    // lines per lane (x 128 per quad in mode 2) = 2^25 lines; one untimed
    // launch first
    uint32_t *sink = (uint32_t *) (buf + nlines * 64);
    const OccLine *src = (const OccLine *) buf;
    auto launch = [&]() {
        const dim3 g(4096), b(256);
        switch (mode) {
        case 0:
            hipLaunchKernelGGL(k_gather_probe<0>, g, b, 0, ctx->stream, src, nlines, 32u, sink);
            break;
        case 1:
            hipLaunchKernelGGL(k_gather_probe<1>, g, b, 0, ctx->stream, src, nlines, 32u, sink);
            break;
        case 2:
            hipLaunchKernelGGL(k_gather_probe<2>, g, b, 0, ctx->stream, src, nlines, 128u, sink);
            break;
        default:
            hipLaunchKernelGGL(k_gather_probe<3>, g, b, 0, ctx->stream, src, nlines, 32u, sink);
        }
    };
    launch();
    hipEvent_t a, b;
    PQ_CHECK(hipEventCreate(&a));
    PQ_CHECK(hipEventCreate(&b));
    PQ_CHECK(hipEventRecord(a, ctx->stream));
    launch();
    PQ_CHECK(hipEventRecord(b, ctx->stream));
    PQ_CHECK(hipEventSynchronize(b));
    float t = 0;
    (void) hipEventElapsedTime(&t, a, b);
    if (ms)
        *ms = t;
    (void) hipEventDestroy(a);
    (void) hipEventDestroy(b);
    return PQ_OK;
}

extern "C" int pq_get_timing(pq_ctx_t *ctx, pq_timing_t *t)
{
    if (!ctx || !t)
        return PQ_ERR_ARG;
    *t = ctx->timing;
    return PQ_OK;
}
