// pq_dedup.hip -- duplicate removal of the reads import on the device.
//
// import_singlereads (index.c:327-333) drops a read when bloom_filter
// (bloom.c:91-135) reports it seen: all of the read's n_hashfunc filter bits
// are set, or, failing that, all of its reverse complement's bits are; a read
// that is kept sets its own (forward) bits.  Reads with N take no part.  Hashes
// are bloom.c:18-86 over the code string (codes 1..4) in the index width.
//
// The verdict of read i depends on the bits set by the kept reads before it, a
// chain through the whole file.  Here it is resolved exactly and in parallel:
//
//   keep_i = (F_i not within S_i) and (R_i not within S_i),
//   S_i    = union of F_j over kept j < i,
//
// with F_i / R_i the bit positions of the read / its reverse complement.  With
// first[b] = the first kept read whose F holds b, "b is not in S_i" is
// first[b] >= i.  Starting from every read kept, each pass recomputes first[]
// from the current verdicts and then every verdict from first[] (a Jacobi
// sweep of a system in which read i depends on reads before i only).  After
// pass k the verdicts of reads 0..k-1 are final, so the sweep ends with no
// change after at most n + 1 passes, and a fixed point is the unique solution:
// the reference's sequential outcome.  Real read sets settle in a few passes
// (a duplicate's verdict depends on the first copy, whose own verdict rarely
// depends on anything).
#include <algorithm>
#include <mutex>

#include "pq_ctx.h"

namespace {

constexpr uint32_t FIRST_NONE = 0xFFFFFFFFu;

// bloom.c:18-86, all five over one pass of the string; h % size at the end.
template <typename T> struct Hash5 {
    T h0 = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0;
    __device__ __forceinline__ void step(T c)
    {
        h0 ^= (T) ((h0 << 5) + (h0 >> 2) + c);
        h1 = (T) (c + (h1 << 6) + (h1 << 16) - h1);
        h2 = (T) (h2 + c);
        h3 ^= c;
        h4 = (T) (33 * h4 + c);
    }
    __device__ __forceinline__ void store(uint64_t *out, uint32_t nf, T size) const
    {
        out[0] = (uint64_t) (h0 % size);
        if (nf > 1)
            out[1] = (uint64_t) (h1 % size);
        if (nf > 2)
            out[2] = (uint64_t) (h2 % size);
        if (nf > 3)
            out[3] = (uint64_t) (h3 % size);
        if (nf > 4)
            out[4] = (uint64_t) (h4 % size);
    }
};

// Filter bit positions of every read (hf) and of its reverse complement (hr):
// one lane per read, nf per orientation.
template <typename T>
__global__ void k_read_hashes(const uint8_t *codes, const uint8_t *ok, uint64_t nrec, uint32_t L,
                              T size, uint32_t nf, uint64_t *hf, uint64_t *hr)
{
    PQ_GRID_STRIDE_FOR(r, nrec)
    {
        if (!ok[r])
            continue;
        const uint8_t *s = codes + r * L;
        Hash5<T> f, b;
        for (uint32_t j = 0; j < L; j++) {
            f.step((T) s[j]);
            b.step((T) (5u - s[L - 1 - j]));   // complement of the code read backwards
        }
        f.store(hf + r * nf, nf, size);
        b.store(hr + r * nf, nf, size);
    }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t count, uint32_t v)
{
    PQ_GRID_STRIDE_FOR(i, count)
        p[i] = v;
}

__global__ void k_keep_init(const uint8_t *ok, uint64_t nrec, uint8_t *keep, uint32_t *changed)
{
    PQ_GRID_STRIDE_FOR(r, nrec)
        keep[r] = ok[r];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *changed = 0;
}

// Clears the first[] entries the previous pass may have set (the forward bits
// of every read that takes part).
__global__ void k_first_clear(const uint8_t *ok, const uint64_t *hf, uint64_t nrec, uint32_t nf,
                              uint32_t *first, uint32_t *changed)
{
    PQ_GRID_STRIDE_FOR(r, nrec)
    {
        if (!ok[r])
            continue;
        for (uint32_t i = 0; i < nf; i++)
            first[hf[r * nf + i]] = FIRST_NONE;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *changed = 0;
}

__global__ void k_first_mark(const uint8_t *keep, const uint64_t *hf, uint64_t nrec, uint32_t nf,
                             uint32_t *first)
{
    PQ_GRID_STRIDE_FOR(r, nrec)
    {
        if (!keep[r])
            continue;
        for (uint32_t i = 0; i < nf; i++)
            atomicMin(first + hf[r * nf + i], (uint32_t) r);
    }
}

__global__ void k_verdicts(const uint8_t *ok, const uint64_t *hf, const uint64_t *hr,
                           const uint32_t *first, uint64_t nrec, uint32_t nf, uint8_t *keep,
                           uint32_t *changed)
{
    PQ_GRID_STRIDE_FOR(r, nrec)
    {
        if (!ok[r])
            continue;
        bool fnew = false, rnew = false;
        for (uint32_t i = 0; i < nf; i++) {
            fnew |= first[hf[r * nf + i]] >= (uint32_t) r;
            rnew |= first[hr[r * nf + i]] >= (uint32_t) r;
        }
        const uint8_t k = fnew && rnew ? 1 : 0;
        if (k != keep[r]) {
            keep[r] = k;
            *changed = 1;
        }
    }
}

// The passes over records already on the device: hashes, then first[] and the
// verdicts until no verdict changes.  hf/hr: nrec * nf positions each, first:
// bits entries, changed: one word; every one of them is filled here before it is
// read, so the buffers may hold anything on entry.
template <typename T>
int dedup_passes(pq_ctx_t *ctx, const uint8_t *d_codes, const uint8_t *d_ok, uint64_t nrec,
                 uint32_t L, uint64_t bits, uint32_t nf, uint8_t *d_keep, uint64_t *d_hf,
                 uint64_t *d_hr, uint32_t *d_first, uint32_t *d_changed, uint32_t *passes)
{
    hipStream_t s = ctx->stream;
    const unsigned B = 256, gr = pq_blocks_gs(nrec, B);
    hipLaunchKernelGGL(k_read_hashes<T>, dim3(gr), dim3(B), 0, s, d_codes, d_ok, nrec, L,
                       (T) bits, nf, d_hf, d_hr);
    hipLaunchKernelGGL(k_fill_u32, dim3(pq_blocks_gs(bits, B)), dim3(B), 0, s, d_first, bits,
                       FIRST_NONE);
    hipLaunchKernelGGL(k_keep_init, dim3(gr), dim3(B), 0, s, d_ok, nrec, d_keep, d_changed);
    PQ_CHECK(hipGetLastError());
    // after pass k the first k verdicts are final: n + 1 passes always suffice
    for (uint64_t pass = 1;; pass++) {
        if (pass > 1)
            hipLaunchKernelGGL(k_first_clear, dim3(gr), dim3(B), 0, s, d_ok, d_hf, nrec, nf,
                               d_first, d_changed);
        hipLaunchKernelGGL(k_first_mark, dim3(gr), dim3(B), 0, s, d_keep, d_hf, nrec, nf,
                           d_first);
        hipLaunchKernelGGL(k_verdicts, dim3(gr), dim3(B), 0, s, d_ok, d_hf, d_hr, d_first, nrec,
                           nf, d_keep, d_changed);
        PQ_CHECK(hipGetLastError());
        uint32_t changed = 0;
        PQ_CHECK(hipMemcpyAsync(&changed, d_changed, 4, hipMemcpyDeviceToHost, s));
        PQ_CHECK(hipStreamSynchronize(s));
        if (!changed) {
            if (passes)
                *passes = (uint32_t) std::min<uint64_t>(pass, 0xFFFFFFFFu);
            return PQ_OK;
        }
        if (pass > nrec + 1)
            return ctx->fail(PQ_ERR_STATE, "duplicate removal did not settle");
    }
}

template <typename T>
int dedup_t(pq_ctx_t *ctx, const uint8_t *codes, const uint8_t *ok, uint64_t nrec, uint32_t L,
            uint64_t bits, uint32_t nf, uint8_t *keep, uint32_t *passes)
{
    uint8_t *d_codes = nullptr, *d_ok = nullptr, *d_keep = nullptr;
    uint64_t *d_hf = nullptr, *d_hr = nullptr;
    uint32_t *d_first = nullptr, *d_changed = nullptr;
    auto release = [&]() {
        pq_free(d_codes);
        pq_free(d_ok);
        pq_free(d_keep);
        pq_free(d_hf);
        pq_free(d_hr);
        pq_free(d_first);
        pq_free(d_changed);
    };
    int rc;
    if ((rc = pq_alloc(ctx, (void **) &d_codes, nrec * L)) ||
        (rc = pq_alloc(ctx, (void **) &d_ok, nrec)) ||
        (rc = pq_alloc(ctx, (void **) &d_keep, nrec)) ||
        (rc = pq_alloc(ctx, (void **) &d_hf, nrec * nf * 8)) ||
        (rc = pq_alloc(ctx, (void **) &d_hr, nrec * nf * 8)) ||
        (rc = pq_alloc(ctx, (void **) &d_first, bits * 4)) ||
        (rc = pq_alloc(ctx, (void **) &d_changed, 4))) {
        release();
        return rc;
    }
    hipStream_t s = ctx->stream;
    auto run = [&]() -> int {
        PQ_CHECK(hipMemcpyAsync(d_codes, codes, nrec * L, hipMemcpyHostToDevice, s));
        PQ_CHECK(hipMemcpyAsync(d_ok, ok, nrec, hipMemcpyHostToDevice, s));
        int rc = dedup_passes<T>(ctx, d_codes, d_ok, nrec, L, bits, nf, d_keep, d_hf, d_hr,
                                 d_first, d_changed, passes);
        if (rc)
            return rc;
        PQ_CHECK(hipMemcpyAsync(keep, d_keep, nrec, hipMemcpyDeviceToHost, s));
        PQ_CHECK(hipStreamSynchronize(s));
        return PQ_OK;
    };
    rc = run();
    release();
    return rc;
}

}  // namespace

// Scratch of pq_dedup_reads_noalloc: one copy per device, held until the
// process ends.  In host code the name `scratch` is a host-side stand-in, not a
// device address: host code uses only the per-device address that
// hipGetSymbolAddress returns (noalloc_scratch).
static __device__ __align__(256) uint8_t scratch[PQ_NOALLOC_SCRATCH_BYTES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static void *noalloc_ptr[NOALLOC_MAX_DEVICES];

// Address of this device's copy of `scratch`, looked up on the first call on
// the device (the device current), or nullptr.
static void *noalloc_scratch(pq_ctx_t *ctx)
{
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return nullptr;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[ctx->device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void) hipGetLastError();
            return nullptr;
        }
        noalloc_ptr[ctx->device] = p;
    }
    return noalloc_ptr[ctx->device];
}


extern "C" int pq_dedup_reads(pq_ctx_t *ctx, const uint8_t *codes, const uint8_t *ok,
                              uint64_t nrec, int read_len, uint64_t bloom_bits, int n_hashfunc,
                              int index_bits, uint8_t *keep, uint32_t *passes)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (passes)
        *passes = 0;
    if (read_len <= 0 || n_hashfunc < 1 || n_hashfunc > 5 ||
        (index_bits != 32 && index_bits != 64) || (nrec && (!codes || !ok || !keep)))
        return ctx->fail(PQ_ERR_ARG, "pq_dedup_reads: bad arguments");
    if (nrec == 0)
        return PQ_OK;
    if (bloom_bits == 0) {
        // the reference's h % 0 faults at the first read that reaches the filter
        for (uint64_t r = 0; r < nrec; r++)
            if (ok[r])
                return ctx->fail(PQ_ERR_LIMIT, "duplicate filter of 0 bits");
        std::fill(keep, keep + nrec, (uint8_t) 0);
        return PQ_OK;
    }
    if (index_bits == 32 && bloom_bits > 0xFFFFFFFFull)
        return ctx->fail(PQ_ERR_LIMIT, "duplicate filter beyond the 32-bit index width");
    if (nrec >= FIRST_NONE)
        return ctx->fail(PQ_ERR_LIMIT, "duplicate removal over 2^32 - 1 records or more");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != ctx->device)
        if (hipSetDevice(ctx->device) != hipSuccess)
            return ctx->fail(PQ_ERR_HIP, "hipSetDevice");
    return index_bits == 32
               ? dedup_t<uint32_t>(ctx, codes, ok, nrec, (uint32_t) read_len, bloom_bits,
                                   (uint32_t) n_hashfunc, keep, passes)
               : dedup_t<uint64_t>(ctx, codes, ok, nrec, (uint32_t) read_len, bloom_bits,
                                   (uint32_t) n_hashfunc, keep, passes);
}

static uint64_t noalloc_bytes(uint64_t nrec, uint64_t bloom_bits, int n_hashfunc)
{
    // hash positions of the reads and their reverse complements, first[] per
    // filter bit, the change flag; each part on a 256-byte boundary
    auto up = [](uint64_t b) { return (b + 255) & ~(uint64_t) 255; };
    return 2 * up(nrec * (uint64_t) n_hashfunc * 8) + up(bloom_bits * 4) + up(4);
}

extern "C" int pq_dedup_reads_noalloc(pq_ctx_t *ctx, const uint8_t *d_codes, const uint8_t *d_ok,
                                      uint64_t nrec, int read_len, uint64_t bloom_bits,
                                      int n_hashfunc, int index_bits, uint8_t *d_keep,
                                      uint32_t *passes)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (passes)
        *passes = 0;
    if (read_len <= 0 || n_hashfunc < 1 || n_hashfunc > 5 ||
        (index_bits != 32 && index_bits != 64) || (nrec && (!d_codes || !d_ok || !d_keep)))
        return ctx->fail(PQ_ERR_ARG, "pq_dedup_reads_noalloc: bad arguments");
    if (nrec == 0)
        return PQ_OK;
    if (bloom_bits == 0)
        return ctx->fail(PQ_ERR_LIMIT, "duplicate filter of 0 bits");
    if (index_bits == 32 && bloom_bits > 0xFFFFFFFFull)
        return ctx->fail(PQ_ERR_LIMIT, "duplicate filter beyond the 32-bit index width");
    const uint64_t need = noalloc_bytes(nrec, bloom_bits, n_hashfunc);
    if (nrec >= FIRST_NONE || bloom_bits > PQ_NOALLOC_SCRATCH_BYTES ||
        need > PQ_NOALLOC_SCRATCH_BYTES)
        return ctx->fail(PQ_ERR_LIMIT,
                         "pq_dedup_reads_noalloc: " + std::to_string(nrec) + " records and a filter of " +
                             std::to_string(bloom_bits) + " bits need " + std::to_string(need) +
                             " bytes of scratch; its fixed scratch holds " +
                             std::to_string((uint64_t) PQ_NOALLOC_SCRATCH_BYTES));
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != ctx->device)
        if (hipSetDevice(ctx->device) != hipSuccess)
            return ctx->fail(PQ_ERR_HIP, "hipSetDevice");
    uint8_t *base = (uint8_t *) noalloc_scratch(ctx);
    if (!base)
        return ctx->fail(PQ_ERR_HIP, "address of the fixed dedup scratch");
    auto up = [](uint64_t b) { return (b + 255) & ~(uint64_t) 255; };
    const uint32_t nf = (uint32_t) n_hashfunc;
    uint64_t *d_hf = (uint64_t *) base;
    uint64_t *d_hr = (uint64_t *) (base + up(nrec * nf * 8));
    uint32_t *d_first = (uint32_t *) (base + 2 * up(nrec * nf * 8));
    uint32_t *d_changed = (uint32_t *) (base + 2 * up(nrec * nf * 8) + up(bloom_bits * 4));
    return index_bits == 32
               ? dedup_passes<uint32_t>(ctx, d_codes, d_ok, nrec, (uint32_t) read_len, bloom_bits,
                                        nf, d_keep, d_hf, d_hr, d_first, d_changed, passes)
               : dedup_passes<uint64_t>(ctx, d_codes, d_ok, nrec, (uint32_t) read_len, bloom_bits,
                                        nf, d_keep, d_hf, d_hr, d_first, d_changed, passes);
}
