// pq_prims.hip -- device-wide primitives written for wave64 / gfx950:
// exclusive scan (reduce-then-scan, one wave per 1024-element strip) and the
// array forms of the stable LSD radix pass of pq_radix.h over (uint32 or uint64
// key, uint32 or uint64 value) pairs.
#include "pq_ctx.h"
#include "pq_radix.h"

#include <cstring>

#define SCAN_ITEMS 16
#define SCAN_THREADS 256
#define SCAN_TILE (SCAN_ITEMS * SCAN_THREADS)

struct U4 {
    uint32_t x, y, z, w;
};
__host__ __device__ inline U4 operator+(const U4 &a, const U4 &b)
{
    return U4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
}
template <typename T> __device__ inline T zero_of();
template <> __device__ inline uint32_t zero_of<uint32_t>() { return 0u; }
template <> __device__ inline uint64_t zero_of<uint64_t>() { return 0ull; }
template <> __device__ inline U4 zero_of<U4>() { return U4{0, 0, 0, 0}; }

__device__ inline uint32_t shfl_up_t(uint32_t v, int d) { return __shfl_up(v, d, 64); }
__device__ inline uint64_t shfl_up_t(uint64_t v, int d) { return __shfl_up(v, d, 64); }
__device__ inline U4 shfl_up_t(U4 v, int d)
{
    return U4{__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64), __shfl_up(v.z, d, 64),
              __shfl_up(v.w, d, 64)};
}
__device__ inline uint32_t shfl_t(uint32_t v, int s) { return __shfl(v, s, 64); }
__device__ inline uint64_t shfl_t(uint64_t v, int s) { return __shfl(v, s, 64); }
__device__ inline U4 shfl_t(U4 v, int s)
{
    return U4{__shfl(v.x, s, 64), __shfl(v.y, s, 64), __shfl(v.z, s, 64), __shfl(v.w, s, 64)};
}

// uint subtraction helper for the exclusive value (inclusive minus own value), per type
__device__ inline uint32_t sub_t(uint32_t a, uint32_t b) { return a - b; }
__device__ inline uint64_t sub_t(uint64_t a, uint64_t b) { return a - b; }
__device__ inline U4 sub_t(U4 a, U4 b) { return U4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }

template <typename T> __device__ inline T wave_incl_scan(T v, int lane)
{
    for (int d = 1; d < 64; d <<= 1) {
        T o = shfl_up_t(v, d);
        if (lane >= d)
            v = v + o;
    }
    return v;
}

// Pass 1: per-tile sums.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const T *in, uint64_t count, T *sums)
{
    __shared__ T ws[SCAN_THREADS / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t base = (uint64_t) blockIdx.x * SCAN_TILE + (uint64_t) wid * (SCAN_TILE / 4);
    T acc = zero_of<T>();
    for (int r = 0; r < SCAN_ITEMS; r++) {
        uint64_t i = base + (uint64_t) r * 64 + lane;
        if (i < count)
            acc = acc + in[i];
    }
    acc = wave_incl_scan(acc, lane);
    if (lane == 63)
        ws[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = zero_of<T>();
        for (int w = 0; w < SCAN_THREADS / 64; w++)
            s = s + ws[w];
        sums[blockIdx.x] = s;
    }
}

// Pass 2: exclusive scan of the tile sums by one block; total to *total.
template <typename T>
__global__ void __launch_bounds__(1024) k_scan_sums(T *sums, uint64_t count, T *total)
{
    __shared__ T ws[16];
    __shared__ T carry_s;
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0)
        carry_s = zero_of<T>();
    __syncthreads();
    for (uint64_t base = 0; base < count; base += 1024) {
        uint64_t i = base + threadIdx.x;
        T v = i < count ? sums[i] : zero_of<T>();
        T inc = wave_incl_scan(v, lane);
        if (lane == 63)
            ws[wid] = inc;
        __syncthreads();
        T wpre = zero_of<T>();
        for (int w = 0; w < wid; w++)
            wpre = wpre + ws[w];
        T c = carry_s;
        if (i < count)
            sums[i] = c + wpre + sub_t(inc, v);
        __syncthreads();
        if (threadIdx.x == 1023) {
            T all = zero_of<T>();
            for (int w = 0; w < 16; w++)
                all = all + ws[w];
            carry_s = c + all;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && total)
        *total = carry_s;
}

// Pass 3: exclusive scan of every tile with its offset.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(const T *in, T *out, uint64_t count,
                                                            const T *sums)
{
    __shared__ T ws[SCAN_THREADS / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t base = (uint64_t) blockIdx.x * SCAN_TILE + (uint64_t) wid * (SCAN_TILE / 4);
    // wave totals first, to get each wave's offset inside the tile
    T acc = zero_of<T>();
    for (int r = 0; r < SCAN_ITEMS; r++) {
        uint64_t i = base + (uint64_t) r * 64 + lane;
        if (i < count)
            acc = acc + in[i];
    }
    acc = wave_incl_scan(acc, lane);
    if (lane == 63)
        ws[wid] = acc;
    __syncthreads();
    T carry = sums[blockIdx.x];
    for (int w = 0; w < wid; w++)
        carry = carry + ws[w];
    for (int r = 0; r < SCAN_ITEMS; r++) {
        uint64_t i = base + (uint64_t) r * 64 + lane;
        T v = i < count ? in[i] : zero_of<T>();
        T inc = wave_incl_scan(v, lane);
        if (i < count)
            out[i] = carry + sub_t(inc, v);
        carry = carry + shfl_t(inc, 63);
    }
}

template <typename T>
static int scan_impl(pq_ctx_t *ctx, const T *in, T *out, uint64_t count, T *total_host,
                     int slot)
{
    if (count == 0) {
        if (total_host)
            memset((void *) total_host, 0, sizeof(T));
        return PQ_OK;
    }
    uint64_t nb = (count + SCAN_TILE - 1) / SCAN_TILE;
    T *sums = (T *) ctx->tmp(slot, (nb + 1) * sizeof(T));
    if (!sums)
        return ctx->fail(PQ_ERR_NOMEM, "scan scratch");
    T *tot = sums + nb;
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned) nb), dim3(SCAN_THREADS), 0, ctx->stream,
                       in, count, sums);
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(1024), 0, ctx->stream, sums, nb, tot);
    hipLaunchKernelGGL(k_scan_down<T>, dim3((unsigned) nb), dim3(SCAN_THREADS), 0, ctx->stream,
                       in, out, count, sums);
    PQ_CHECK(hipGetLastError());
    if (total_host) {
        PQ_CHECK(hipMemcpyAsync(total_host, tot, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return PQ_OK;
}

// scratch of a scan of count elements of elem bytes (slot 0)
size_t pq_scan_scratch_bytes(uint64_t count, size_t elem)
{
    return ((count + SCAN_TILE - 1) / SCAN_TILE + 1) * elem;
}

// scratch slot 0 is reserved for scan partial sums
int pq_scan_u32(pq_ctx_t *ctx, const uint32_t *in, uint32_t *out, uint64_t count,
                uint32_t *total_host)
{
    return scan_impl<uint32_t>(ctx, in, out, count, total_host, 0);
}

int pq_scan_u64(pq_ctx_t *ctx, const uint64_t *in, uint64_t *out, uint64_t count,
                uint64_t *total_host)
{
    return scan_impl<uint64_t>(ctx, in, out, count, total_host, 0);
}

int pq_scan_u4(pq_ctx_t *ctx, const void *in, void *out, uint64_t count, void *total_host)
{
    return scan_impl<U4>(ctx, (const U4 *) in, (U4 *) out, count, (U4 *) total_host, 0);
}

// ---- LSD radix pass (kernels in pq_radix.h) ------------------------------------

int pq_radix_pass(pq_ctx_t *ctx, const uint64_t *kin, const uint32_t *vin, uint64_t *kout,
                  uint32_t *vout, uint64_t count, unsigned shift)
{
    return radix_pass_src(ctx, ArraySrc<uint64_t, uint32_t>{kin, vin}, kout, vout, count, shift);
}

int pq_radix_pass(pq_ctx_t *ctx, const uint64_t *kin, const uint64_t *vin, uint64_t *kout,
                  uint64_t *vout, uint64_t count, unsigned shift)
{
    return radix_pass_src(ctx, ArraySrc<uint64_t, uint64_t>{kin, vin}, kout, vout, count, shift);
}

int pq_radix_pass(pq_ctx_t *ctx, const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                  uint32_t *vout, uint64_t count, unsigned shift)
{
    return radix_pass_src(ctx, ArraySrc<uint32_t, uint32_t>{kin, vin}, kout, vout, count, shift);
}

int pq_radix_pass(pq_ctx_t *ctx, const uint32_t *kin, const uint64_t *vin, uint32_t *kout,
                  uint64_t *vout, uint64_t count, unsigned shift)
{
    return radix_pass_src(ctx, ArraySrc<uint32_t, uint64_t>{kin, vin}, kout, vout, count, shift);
}
