// pq_learned.hip -- learned index over the suffix arrays.
//
// Role: the reference starts two of its four backward searches from a
// precomputed "interval shortcut" (sufsort.c:179-189, consumed at
// string_graph.c:203-212, 289-298) and walks the rest from one code.  Here
// every chain of the candidate search (k_find) can start at depth D = min(63, t)
// from intervals found by a learned index instead of D-1 backward steps, each of
// which loads one or two random 64-byte occurrence lines.  Candidates are only
// recorded from depth t on, so no step below D records anything.
//
// Per direction:
//   keys    = the keys of every stride-th suffix-array row (stride 8 by default):
//             the row's first D codes (first code in the top bits of 128), bit 0
//             set when the suffix has at least D codes.  Keys are non-decreasing
//             along the suffix array, so they bracket any pattern's rows to a
//             stride.
//   model   = FP32 MLP on the k-code prefix of a key (k = 21 codes mapped to
//             -1..1) -> 64 ReLU units -> 1 output, fitted to the sampled row's
//             position / number of samples, which routes the key to one of 4096
//             leaves.
//   leaves  = each leaf is piecewise linear in the k-code prefix value x: over
//             the x range of the training keys routed to it, S equal-width
//             pieces (by default about 8 sampled rows each) whose end points are
//             exact (x, first sampled row >= x) pairs.
//   query   = predicted sampled row -> exact lower bound of the pattern's key
//             among the samples by galloping outward from the prediction, then
//             binary search; the rows between two samples are resolved on the
//             suffix array and the text (an outward scan that also finishes
//             narrow chains, or binary searches for wide intervals).  The
//             prediction only narrows the search: the result is exact for any
//             model, including an untrained one.
//
// Training, all on the device with the kernels below: training rows drawn from
// the sampled keys; the output layer first solved in closed form (least squares over the
// hidden activations, Gram matrix accumulated by a device kernel, 65 x 65
// system solved on the host), then every parameter refined by minibatch Adam
// with the forward/backward kernels (block-ordered deterministic reductions).
#include "pq_ctx.h"
#include "pq_chain.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#define MLP_K_MAX 21
#define MLP_H_MAX 64
#define MLP_P_MAX (MLP_H_MAX * MLP_K_MAX + 2 * MLP_H_MAX + 1)
#define TRAIN_BLOCK 64      // samples per block of the gradient kernel (one wave)
#define SEG_MAX 65536       // linear pieces per leaf
#define SEG_AUTO_ROWS 8     // sampled rows per piece when the count is automatic
#define REROUTE_MAX 64      // leaf corrections per query
#define GRAM_SUB 128        // samples staged in LDS per round of the Gram kernel
#define GRAM_CHUNK 1024     // samples per block of the Gram kernel
#define GRAM_DIM (MLP_H_MAX + 2)   // activations, 1, target

// constant address space: wave-uniform reads through it become scalar loads
typedef __attribute__((address_space(4))) const float cfloat;

struct MlpDims {
    uint32_t k, h;
};

// parameter vector layout: W1[h][k], b1[h], w2[h], b2
static __host__ __device__ inline uint32_t mlp_nparams(uint32_t k, uint32_t h)
{
    return h * k + h + h + 1;
}

// search statistics per index direction: lookups, key probes, sum and max of
// |exact entry - predicted entry|; patterns absent from the table
struct QStats {
    unsigned long long lookups[2], probes[2], win_sum[2], win_max[2], absent;
};
// one copy per WSTAT_SLOTS (blocks add to copy blockIdx % WSTAT_SLOTS)
#define QSTATS_BYTES ((size_t) WSTAT_SLOTS * WSTAT_STRIDE * 8)

struct DirModel {
    u128 *key = nullptr;          // keys of every stride-th SA row (a scratch slot of ctx)
    uint64_t nU = 0;              // sampled rows
    float *params = nullptr;
    float *iparams = nullptr;     // inference form (rows padded to MLP_K_MAX)
    uint64_t *range = nullptr;    // [xmin, xmax] per leaf
    uint32_t *knots = nullptr;    // (S + 1) entries per leaf
    double loss = 0, fit_loss = 0;
};

struct Learned {
    uint32_t k = 21, h = 64, leaves = 4096, segs = 64, depth = 21, stride = 8;
    DirModel dir[2];
    QStats *qstats = nullptr;     // device
    double train_ms = 0;
};

// ---- forward pass ----------------------------------------------------------------------

__device__ __forceinline__ float mlp_feature(uint64_t x, uint32_t k, uint32_t j)
{
    // code j of the k-code prefix x (first code most significant) mapped to -1..1
    uint32_t c = (uint32_t) (x >> (2 * (k - 1 - j))) & 3u;
    return (float) c * (2.0f / 3.0f) - 1.0f;
}

// Output of the MLP; p may live in LDS or global memory.  With act != nullptr the
// ReLU activations are stored (stride act_stride) for the backward pass.
__device__ __forceinline__ float mlp_forward(const float *p, MlpDims d, uint64_t x, float *act,
                                             uint32_t act_stride)
{
    float f[MLP_K_MAX];
#pragma unroll
    for (uint32_t j = 0; j < MLP_K_MAX; j++)
        f[j] = j < d.k ? mlp_feature(x, d.k, j) : 0.0f;
    const float *W1 = p, *b1 = p + d.h * d.k, *w2 = b1 + d.h;
    float out = w2[d.h];   // b2
    for (uint32_t i = 0; i < d.h; i++) {
        float s = b1[i];
        const float *row = W1 + i * d.k;
#pragma unroll
        for (uint32_t j = 0; j < MLP_K_MAX; j++)
            if (j < d.k)
                s = fmaf(row[j], f[j], s);
        float a = s > 0.0f ? s : 0.0f;
        if (act)
            act[i * act_stride] = a;
        out = fmaf(w2[i], a, out);
    }
    return out;
}

// Inference form used by the search: W1 rows padded to MLP_K_MAX columns with
// zero weights (features past k are 0 too), then b1, w2, b2.  Fixed trip counts
// and wave-uniform parameters let every weight come from a scalar load.  Gives
// the same output as mlp_forward on the unpadded parameters.
template <typename PT>
__device__ __forceinline__ float mlp_forward_padded(const PT *p, MlpDims d, uint64_t x)
{
    float f[MLP_K_MAX];
#pragma unroll
    for (uint32_t j = 0; j < MLP_K_MAX; j++) {
        uint32_t sh = j < d.k ? 2 * (d.k - 1 - j) : 0;
        float v = (float) ((uint32_t) (x >> sh) & 3u) * (2.0f / 3.0f) - 1.0f;
        f[j] = j < d.k ? v : 0.0f;
    }
    const PT *b1 = p + d.h * MLP_K_MAX, *w2 = b1 + d.h;
    float out = w2[d.h];
    for (uint32_t i = 0; i < d.h; i++) {
        const PT *row = p + i * MLP_K_MAX;
        float s = b1[i];
#pragma unroll
        for (uint32_t j = 0; j < MLP_K_MAX; j++)
            s = fmaf(row[j], f[j], s);
        float a = s > 0.0f ? s : 0.0f;
        out = fmaf(w2[i], a, out);
    }
    return out;
}

// ---- training ----------------------------------------------------------------------

// One wave per block, one sample per lane: forward, backward, and the block's
// summed gradient of inv_batch * sum (out - target)^2, written to part[block][P];
// the block's summed squared residual goes to loss_part[block].
__global__ void __launch_bounds__(TRAIN_BLOCK) k_mlp_grad(const float *params, MlpDims d,
                                                          const uint64_t *xs,
                                                          const float *targets, uint32_t n,
                                                          float inv_batch, float *part,
                                                          float *loss_part)
{
    __shared__ float sp[MLP_P_MAX];
    __shared__ float sf[TRAIN_BLOCK][MLP_K_MAX + 1];
    __shared__ float sa[TRAIN_BLOCK][MLP_H_MAX + 1];   // ReLU activations
    __shared__ float sd[TRAIN_BLOCK][MLP_H_MAX + 1];   // dLoss/dpre
    __shared__ float se[TRAIN_BLOCK];                  // dLoss/dout
    __shared__ float sl[TRAIN_BLOCK];                  // squared residual
    const uint32_t P = mlp_nparams(d.k, d.h);
    const uint32_t tx = threadIdx.x;
    for (uint32_t i = tx; i < P; i += blockDim.x)
        sp[i] = params[i];
    __syncthreads();
    const uint32_t s = blockIdx.x * TRAIN_BLOCK + tx;
    const bool ok = s < n;
    const uint64_t x = ok ? xs[s] : 0;
    float out = mlp_forward(sp, d, x, &sa[tx][0], 1);
    float r = ok ? out - targets[s] : 0.0f;
    float e = 2.0f * r * inv_batch;
    se[tx] = e;
    sl[tx] = r * r;
    for (uint32_t j = 0; j < d.k; j++)
        sf[tx][j] = mlp_feature(x, d.k, j);
    const float *w2 = sp + d.h * d.k + d.h;
    for (uint32_t i = 0; i < d.h; i++)
        sd[tx][i] = sa[tx][i] > 0.0f ? e * w2[i] : 0.0f;
    __syncthreads();
    // dW1[i][j] = sum_t dpre_i f_j; db1[i] = sum_t dpre_i; dw2[i] = sum_t e a_i; db2 = sum_t e
    float *o = part + (uint64_t) blockIdx.x * P;
    const uint32_t nw = d.h * d.k;
    for (uint32_t q = tx; q < P; q += blockDim.x) {
        float g = 0.0f;
        if (q < nw) {
            uint32_t i = q / d.k, j = q - i * d.k;
            for (uint32_t t = 0; t < TRAIN_BLOCK; t++)
                g = fmaf(sd[t][i], sf[t][j], g);
        } else if (q < nw + d.h) {
            uint32_t i = q - nw;
            for (uint32_t t = 0; t < TRAIN_BLOCK; t++)
                g += sd[t][i];
        } else if (q < nw + 2 * d.h) {
            uint32_t i = q - nw - d.h;
            for (uint32_t t = 0; t < TRAIN_BLOCK; t++)
                g = fmaf(se[t], sa[t][i], g);
        } else {
            for (uint32_t t = 0; t < TRAIN_BLOCK; t++)
                g += se[t];
        }
        o[q] = g;
    }
    if (tx == 0) {
        float acc = 0.0f;
        for (uint32_t t = 0; t < TRAIN_BLOCK; t++)
            acc += sl[t];
        loss_part[blockIdx.x] = acc;
    }
}

// Block partials summed in block order (deterministic).
template <typename T>
__global__ void k_sum_parts(const T *part, uint32_t nblocks, uint32_t P, T *sum)
{
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P)
        return;
    T g = 0;
    for (uint32_t b = 0; b < nblocks; b++)
        g += part[(uint64_t) b * P + q];
    sum[q] = g;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the block-summed gradient.
__global__ void k_mlp_adam(float *params, float *m, float *v, const float *part, uint32_t nblocks,
                           uint32_t P, float lr, float b1t, float b2t)
{
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P)
        return;
    float g = 0.0f;
    for (uint32_t b = 0; b < nblocks; b++)
        g += part[(uint64_t) b * P + q];
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float mq = beta1 * m[q] + (1.0f - beta1) * g;
    float vq = beta2 * v[q] + (1.0f - beta2) * g * g;
    m[q] = mq;
    v[q] = vq;
    float mh = mq / (1.0f - b1t), vh = vq / (1.0f - b2t);
    params[q] -= lr * mh / (sqrtf(vh) + eps);
}

// Upper triangle (in double) of the Gram matrix of z = (a_0 .. a_{h-1}, 1, y)
// over a block's GRAM_CHUNK samples, a = hidden activations of the forward pass:
// the normal equations of the least-squares output layer.
__global__ void __launch_bounds__(256) k_ls_gram(const float *params, MlpDims d,
                                                 const uint64_t *xs, const float *targets,
                                                 uint32_t n, double *part)
{
    __shared__ float z[GRAM_SUB][GRAM_DIM + 1];
    __shared__ uint8_t pi[GRAM_DIM * (GRAM_DIM + 1) / 2], pj[GRAM_DIM * (GRAM_DIM + 1) / 2];
    const uint32_t m = d.h + 2, ne = m * (m + 1) / 2;
    if (threadIdx.x == 0) {
        uint32_t e = 0;
        for (uint32_t i = 0; i < m; i++)
            for (uint32_t j = i; j < m; j++, e++) {
                pi[e] = (uint8_t) i;
                pj[e] = (uint8_t) j;
            }
    }
    constexpr uint32_t per = (GRAM_DIM * (GRAM_DIM + 1) / 2 + 255) / 256;
    double acc[per];
    for (uint32_t q = 0; q < per; q++)
        acc[q] = 0.0;
    const uint64_t base = (uint64_t) blockIdx.x * GRAM_CHUNK;
    for (uint32_t sub = 0; sub < GRAM_CHUNK; sub += GRAM_SUB) {
        __syncthreads();
        if (threadIdx.x < GRAM_SUB) {
            uint64_t s = base + sub + threadIdx.x;
            float *zr = &z[threadIdx.x][0];
            if (s < n) {
                mlp_forward(params, d, xs[s], zr, 1);
                zr[d.h] = 1.0f;
                zr[d.h + 1] = targets[s];
            } else {
                for (uint32_t i = 0; i < m; i++)
                    zr[i] = 0.0f;
            }
        }
        __syncthreads();
        for (uint32_t q = 0; q < per; q++) {
            uint32_t e = threadIdx.x + q * 256;
            if (e < ne) {
                uint32_t i = pi[e], j = pj[e];
                double a = 0.0;
                for (uint32_t t = 0; t < GRAM_SUB; t++)
                    a += (double) z[t][i] * (double) z[t][j];
                acc[q] += a;
            }
        }
    }
    for (uint32_t q = 0; q < per; q++) {
        uint32_t e = threadIdx.x + q * 256;
        if (e < ne)
            part[(uint64_t) blockIdx.x * ne + e] = acc[q];
    }
}

// ---- keys, table, samples, leaves ------------------------------------------------------

// The exact part of the learned index: K[j] = key of suffix-array row j * stride,
// j < ns (every stride-th row; the suffix array is read once, the text at one
// position per stride rows).
template <typename R>
__global__ void k_sample_keys(Geom G, const uint64_t *t2, const R *sa, uint32_t stride,
                              uint64_t ns, uint32_t D, u128 *K)
{
    PQ_GRID_STRIDE_FOR(j, ns)
        K[j] = suffix_key(G, t2, (uint64_t) sa[j * stride], D);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Training samples drawn uniformly from the sampled rows: the k-code prefix of
// the key and (sampled row) / nU, the suffix-array row / N to the stride.
__global__ void k_sample_rows(const u128 *key, uint64_t nU, uint32_t k, uint32_t n, uint64_t seed,
                              uint64_t *xs, float *targets)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    uint64_t u = (uint64_t) (((u128) mix64(seed + 0x9E3779B97F4A7C15ull * (s + 1)) * nU) >> 64);
    xs[s] = (uint64_t) (key[u] >> (128 - 2 * k));
    targets[s] = (float) ((double) u / (double) nU);
}

__device__ __forceinline__ uint32_t leaf_of(float y, uint32_t leaves)
{
    float z = floorf(y * (float) leaves);
    if (!(z >= 0.0f))
        return 0;
    return z >= (float) leaves ? leaves - 1 : (uint32_t) z;
}

__global__ void k_mlp_route(const float *params, MlpDims d, const uint64_t *xs, uint32_t n,
                            uint32_t leaves, uint32_t *leaf)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    leaf[s] = leaf_of(mlp_forward(params, d, xs[s], nullptr, 0), leaves);
}

// First entry u with key[u] >= q (plain binary search over [0, n)).
__device__ __forceinline__ uint64_t lower_bound_all(const u128 *key, uint64_t n, u128 q)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (key[mid] < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t knot_x(uint64_t xmin, uint64_t xmax, uint32_t S, uint32_t s)
{
    return xmin + ((xmax - xmin) * s) / S;
}

// Exact end points of every piece: the first sampled row at or after the knot's
// k-code prefix.  One block per leaf: the knots of leaf j lie between the lower
// bounds of its range's two ends, which two lanes find over all keys; every knot
// is then searched within that bracket only.
__global__ void __launch_bounds__(256) k_leaf_knots(const u128 *key, uint64_t nU, uint32_t k,
                                                    const uint64_t *range, uint32_t S,
                                                    uint32_t *knots)
{
    const uint32_t j = blockIdx.x;
    const uint64_t xmin = range[2 * j], xmax = range[2 * j + 1];
    const unsigned sh = 128 - 2 * k;
    __shared__ uint64_t br[2];
    if (threadIdx.x < 2)
        br[threadIdx.x] = lower_bound_all(key, nU, (u128) (threadIdx.x ? xmax : xmin) << sh);
    __syncthreads();
    for (uint32_t s = threadIdx.x; s <= S; s += blockDim.x) {
        const u128 q = (u128) knot_x(xmin, xmax, S, s) << sh;
        uint64_t lo = br[0], hi = br[1];   // xmin <= knot <= xmax: its bound is in [lo, hi]
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (key[mid] < q)
                lo = mid + 1;
            else
                hi = mid;
        }
        knots[(uint64_t) j * (S + 1) + s] = (uint32_t) lo;
    }
}

// ---- query ------------------------------------------------------------------------------

template <typename R> struct ModelViewT {
    const u128 *key;              // sampled keys
    uint64_t nU;                  // sampled rows
    uint32_t stride;
    uint64_t N;                   // suffix-array rows
    const R *sa;
    const uint64_t *t2;           // packed text of the direction
    const float *iparams;
    const uint64_t *range;
    const uint32_t *knots;
};

// First row r in [lo, hi) whose suffix key is >= q, given that the key of row
// hi is >= q (or hi = N) and that of row lo - 1 is < q: binary search on the
// suffix array itself (a probe reads the row and the text it points at).
template <typename R>
__device__ __forceinline__ uint64_t row_lower_bound(const Geom &G, const ModelViewT<R> &M,
                                                    uint64_t lo, uint64_t hi, u128 q, uint32_t D,
                                                    uint32_t &probes)
{
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        probes++;
        if (suffix_key(G, M.t2, (uint64_t) M.sa[mid], D) < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Parameters read by every lane of a wave at the same addresses: the pointer is
// made wave-uniform and moved to the constant address space (scalar loads).
__device__ __forceinline__ const cfloat *wave_params(const float *p)
{
    uint64_t v = (uint64_t) p;
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t) v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t) (v >> 32));
    return (const cfloat *) (((uint64_t) hi << 32) | lo);
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Input j of the padded network for prefix value x of k codes: the code mapped
// to -1..1 for j < k, 0 for k <= j < MLP_K_MAX, and 1 for j = MLP_K_MAX (the
// input that carries b1 in the padded weight matrix below).
__device__ __forceinline__ float mlp_input(uint64_t x, uint32_t k, uint32_t j)
{
    if (j == MLP_K_MAX)
        return 1.0f;
    if (j >= k)
        return 0.0f;
    return (float) ((uint32_t) (x >> (2 * (k - 1 - j))) & 3u) * (2.0f / 3.0f) - 1.0f;
}

// Network output for every lane's prefix value x (one key per lane, 64 per
// wave) on the matrix cores, same function as mlp_forward_padded.  With W' =
// [W1 | b1] (h x 22) and X' the 64 x 22 inputs, H^T = W' X'^T (h x 64) is four
// 32 x 32 tiles (hidden half nt, key half mt) accumulated by 11 steps of
// v_mfma_f32_32x32x2f32 each.  Operands per lane l (c = l % 32, g = l / 32):
// A = W'[32 nt + c][k0 + g], B = X'[32 mt + c][k0 + g]; result register v holds
// H^T[32 nt + 8 (v / 4) + 4 g + v % 4][32 mt + c], so a lane holds half the
// hidden units of key 32 mt + c and the lane l ^ 32 the other half.  p: the
// padded parameters (LDS).  Every lane of the wave must call it.
__device__ __forceinline__ float mlp_forward_mfma(const float *p, MlpDims d, uint64_t x)
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5, c = lane & 31;
    const uint64_t xp = __shfl_xor(x, 32, 64);   // the key of lane l ^ 32
    const float *b1 = p + d.h * MLP_K_MAX, *w2 = b1 + d.h;
    float part0 = 0.0f, part1 = 0.0f;
    // one 32 x 32 tile at a time (16 accumulators live), keys half mt, hidden half nt
#pragma unroll 1
    for (int t4 = 0; t4 < 4; t4++) {
        const int mt = t4 >> 1, nt = t4 & 1;
        const uint64_t xb = g == mt ? x : xp;    // key 32 mt + c
        const uint32_t n = 32 * nt + c;
        floatx16 acc = {};
#pragma unroll
        for (uint32_t k0 = 0; k0 < MLP_K_MAX + 1; k0 += 2) {
            const uint32_t kk = k0 + g;
            const float a = n >= d.h ? 0.0f : kk < MLP_K_MAX ? p[n * MLP_K_MAX + kk] : b1[n];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, mlp_input(xb, d.k, kk), acc, 0, 0, 0);
        }
        float part = 0.0f;
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const uint32_t r = 32 * nt + 8 * (v / 4) + 4 * g + (v % 4);
            part = fmaf(r < d.h ? w2[r] : 0.0f, fmaxf(acc[v], 0.0f), part);
        }
        if (mt == 0)
            part0 += part;
        else
            part1 += part;
    }
    part0 += __shfl_xor(part0, 32, 64);
    part1 += __shfl_xor(part1, 32, 64);
    // lane l < 32 is key c = 32 * 0 + c, lane l >= 32 key 32 * 1 + c
    return w2[d.h] + (g == 0 ? part0 : part1);
}

// Predicted entry of prefix value x from the network output y: leaf, then the
// leaf's piecewise-linear piece.
template <typename R>
__device__ __forceinline__ uint64_t leaf_entry(float y, const ModelViewT<R> &M, uint32_t leaves,
                                               uint32_t S, uint64_t x)
{
    uint32_t j = leaf_of(y, leaves);
    // the network routes approximately; step to the neighbouring leaf whose
    // training range holds x (bounded)
    for (int q = 0; q < REROUTE_MAX && j > 0 && x < M.range[2 * j]; q++)
        j--;
    for (int q = 0; q < REROUTE_MAX && j + 1 < leaves && x > M.range[2 * j + 1]; q++)
        j++;
    uint64_t xmin = M.range[2 * j], xmax = M.range[2 * j + 1];
    const uint32_t *kn = M.knots + (uint64_t) j * (S + 1);
    if (x <= xmin)
        return kn[0];
    if (x >= xmax)
        return kn[S];
    uint64_t span = xmax - xmin;
    uint32_t s = (uint32_t) (((x - xmin) * S) / span);
    s = s >= S ? S - 1 : s;
    uint64_t xa = knot_x(xmin, xmax, S, s), xb = knot_x(xmin, xmax, S, s + 1);
    uint32_t ra = kn[s], rb = kn[s + 1];
    if (xb <= xa || x <= xa)
        return ra;
    float frac = (float) (x - xa) / (float) (xb - xa);
    return ra + (uint64_t) (frac * (float) (rb - ra));
}

// First entry u >= 0 with key[u] >= q, searched outward from guess: gallop until
// bracketed, then binary search.  Exact for any guess.
__device__ __forceinline__ uint64_t lower_bound_from(const u128 *key, uint64_t n, u128 q,
                                                     uint64_t guess, uint32_t &probes)
{
    if (n == 0)
        return 0;
    if (guess >= n)
        guess = n - 1;
    uint64_t lo, hi;   // answer in [lo, hi]; key[hi] >= q or hi == n
    probes++;
    if (key[guess] >= q) {
        hi = guess;
        lo = 0;
        for (uint64_t step = 1; hi > 0; step <<= 1) {
            uint64_t c = hi > step ? hi - step : 0;
            probes++;
            if (key[c] < q) {
                lo = c + 1;
                break;
            }
            hi = c;
        }
    } else {
        lo = guess + 1;
        hi = n;
        for (uint64_t step = 1;; step <<= 1) {
            uint64_t c = guess + step;
            if (c >= n)
                break;
            probes++;
            if (key[c] >= q) {
                hi = c;
                break;
            }
            lo = c + 1;
        }
    }
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        probes++;
        if (key[mid] < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t rev_codes64(uint64_t x)
{
    // reverse the order of the 32 two-bit codes of x
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return __builtin_bswap64(x);
}

// The D codes of a left-aligned pattern in reverse order, left-aligned.
__device__ __forceinline__ u128 rev_pattern(u128 p, uint32_t D)
{
    u128 r = ((u128) rev_codes64((uint64_t) p) << 64) | rev_codes64((uint64_t) (p >> 64));
    return r << (128 - 2 * D);
}

// The first D codes a chain has matched at depth D, left-aligned in 128 bits, as
// searched in its index.  Chains as in k_find: FB n[L-D..L) in fw, FF
// comp(n[L-D..L)) in bw, BF rev(n[0..D)) in bw, BB rc(n[0..D)) in fw.
__device__ __forceinline__ u128 chain_pattern(const uint64_t *fwt, uint64_t s, uint32_t L, int ch,
                                              uint32_t D)
{
    const u128 mask = top_mask(D);
    u128 fwd = text_codes(fwt, (ch < 2) ? s + L - D : s) & mask;
    switch (ch) {
    case 0: return fwd;
    case 1: return fwd ^ mask;
    case 2: return rev_pattern(fwd, D);
    default: return rev_pattern(fwd, D) ^ mask;
    }
}

// Interval of a pattern found by scanning the suffix array outward from the
// sampled rows while the chain is finished on the text (one lane).  Row r's key
// (its first D codes, as suffix_key) is read from the other direction's text X,
// where the codes of the suffix at g = start(m) + p stand reversed left of
// xs = start(m) + len(m) - p; the same rows' codes right of xs are the ones
// verify_chain compares with the chain's next codes, so one window per row both
// places the row and, for a member, gives its reach.  m1: row r1 is a member
// (a sampled key equals the pattern's); else the rows in (bl, r1) may hold the
// whole interval, below rows whose key is larger.  Rows below bl and row bl - 1
// have smaller keys.  Every member is processed as in verify_chain (reaches, the
// top two, tentative candidates: at most vmax, staged in stg); a row is visited
// only if its neighbour towards the sampled row belongs to the interval (or lies
// above it), so the scan reads the members and at most one row past either end:
// rows read = size + [lo > bl] + [hi < N] with m1, (r1 - lo) + [lo > bl] without.
// Two rows are visited per iteration (two chains of dependent loads in flight);
// a speculative second row beyond an end changes neither the result nor the
// count of rows read.
struct ScanOut {
    uint64_t lo, hi;         // the pattern's rows [lo, hi)
    uint32_t k;              // candidates kept after the o_last rule (size <= vmax)
    uint32_t visited;        // rows read
};

template <typename R>
__device__ ScanOut scan_chain(const Geom &G, const R *sa, uint64_t N, const uint64_t *X,
                              const uint64_t *Pt, uint64_t flip, uint64_t sn, uint32_t Ln,
                              uint32_t D, u128 key, uint64_t bl, uint64_t r1, bool m1,
                              uint32_t minsize, uint32_t t, uint32_t vmax, uint2 *lc, uint16_t *oo,
                              uint64_t *stg)
{
    const uint32_t o0 = D, jmax = Ln - 1 - o0;
    const uint32_t nw = (jmax + 31) / 32;
    uint64_t pw[8];
    {
        const uint64_t ps = sn + o0, w0 = ps >> 5;
        const unsigned sh = (unsigned) (ps & 31) * 2;
        uint64_t cur = Pt[w0];
#pragma unroll
        for (uint32_t w = 0; w < 8; w++) {
            const uint64_t nxt = w < nw ? Pt[w0 + w + 1] : 0ull;
            pw[w] = ((cur << sh) | ((nxt >> 1) >> (63 - sh))) ^ flip;
            cur = nxt;
        }
    }
    uint32_t t1 = 0, t2 = 0, visited = 0;
    CandStage cs(stg, lc, oo);
    // sign of (key of the row at text position g) - key; a member's reach and
    // tentative candidate are recorded on the way
    auto visit = [&](uint64_t g) -> int {
        const uint64_t m = pq_read_of(G, g);
        const uint64_t sm = pq_read_start(G, m);
        const uint32_t len = pq_read_len(G, m);
        const uint32_t p = (uint32_t) (g - sm), rem = len - p;   // rem >= 1
        const uint64_t xs = sm + rem;
        const uint32_t c = rem < D ? rem : D;
        const u128 rk = rev_pattern(text_codes(X, xs - c), c) | (u128) (rem >= D ? 1u : 0u);
        if (rk != key)
            return rk < key ? -1 : 1;
        const uint32_t J = p < jmax ? p : jmax;
        const uint64_t x0 = xs >> 5;
        const unsigned xsh = (unsigned) (xs & 31) * 2;
        uint64_t cur = X[x0];
        uint32_t e = J;
#pragma unroll
        for (uint32_t w = 0; w < 8; w++) {
            if (32 * w >= J)
                break;
            const uint64_t nxt = X[x0 + w + 1];
            const uint64_t x = pw[w] ^ ((cur << xsh) | ((nxt >> 1) >> (63 - xsh)));
            if (x) {
                const uint32_t f = 32 * w + (uint32_t) __clzll((long long) x) / 2;
                e = f < J ? f : J;
                break;
            }
            cur = nxt;
        }
        const uint32_t v = o0 + e + 1;
        if (v > t1) {
            t2 = t1;
            t1 = v;
        } else if (v > t2) {
            t2 = v;
        }
        if (p <= jmax && e == p && o0 + p >= t && cs.n < vmax)   // tentative
            cs.add((uint32_t) m, o0 + p);
        return 0;
    };
    // rows [lo, hi) are members so far; dn: next row down is dn - 1, up: next row up
    bool found = m1;
    uint64_t lo = r1, hi = r1, dn = r1, up = r1 + 1;
    if (m1) {
        visit((uint64_t) sa[r1]);
        hi = r1 + 1;
        visited = 1;
    }
    bool dact = dn > bl, uact = m1 && up < N;
    while (dact || uact) {
        // two rows: one each way, or two in the one direction still open
        uint64_t rr[2];
        int dir[2];   // -1 down, +1 up, 0 none
        if (dact && uact) {
            rr[0] = dn - 1, dir[0] = -1;
            rr[1] = up, dir[1] = 1;
        } else if (dact) {
            rr[0] = dn - 1, dir[0] = -1;
            rr[1] = dn - 2, dir[1] = dn - 1 > bl ? -1 : 0;
        } else {
            rr[0] = up, dir[0] = 1;
            rr[1] = up + 1, dir[1] = up + 1 < N ? 1 : 0;
        }
        uint64_t g[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
            g[h] = dir[h] ? (uint64_t) sa[rr[h]] : 0ull;
        int cmp[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
            cmp[h] = dir[h] ? visit(g[h]) : 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint64_t r = rr[h];
            // rows whose outcome is used are counted (not a speculative one past an end)
            visited += (dir[h] < 0 && dact) || (dir[h] > 0 && uact);
            if (dir[h] < 0 && dact) {
                if (cmp[h] < 0) {           // below the interval
                    if (!found)
                        lo = hi = r + 1;
                    dact = false;
                } else {
                    if (cmp[h] == 0) {
                        if (!found)
                            hi = r + 1;
                        found = true;
                        lo = r;
                    } else {                // above it: only before the first member
                        lo = hi = r;
                    }
                    dn = r;
                    dact = r > bl;
                }
            } else if (dir[h] > 0 && uact) {
                if (cmp[h] != 0) {
                    uact = false;
                } else {
                    hi = r + 1;
                    up = r + 1;
                    uact = up < N;
                }
            }
        }
    }
    uint32_t k = 0;
    const uint64_t size = hi - lo;
    if (size >= minsize && size <= vmax) {
        const uint32_t top = minsize == 1 ? t1 : t2;   // >= o0 + 1: members reach o0
        k = cs.finish(top - 1 < Ln - 1 ? top - 1 : Ln - 1);
    }
    return {lo, hi, k, visited};
}

// The model's prediction for every chain of a batch that takes a seed: the
// sampled row its depth-D pattern should start at (the network on the matrix
// cores, then the leaf's piecewise-linear piece), written to pred[ch * nr + rl].
// Kept apart from the seed kernel, whose long chain of dependent reads then runs
// without the network's registers and LDS.  Same block shape as the seed kernel:
// wave q handles chain q of 64 reads, so the model of a wave is uniform.
template <typename R>
__global__ void __launch_bounds__(256) k_learned_predict(Geom G, const uint64_t *fwt, uint64_t r0,
                                                         uint64_t nr, uint32_t D, MlpDims d,
                                                         uint32_t leaves, uint32_t S,
                                                         ModelViewT<R> fw, ModelViewT<R> bw,
                                                         uint64_t *pred)
{
    const int ch = (int) (threadIdx.x >> 6);
    const uint64_t rl = (uint64_t) blockIdx.x * 64 + (threadIdx.x & 63);
    const bool fwsearch = (ch == 0 || ch == 3);
    // both directions' inference parameters in LDS, read as wave-wide broadcasts
    __shared__ float wsh[2][MLP_P_MAX];
    for (uint32_t i = threadIdx.x; i < MLP_P_MAX; i += blockDim.x) {
        wsh[0][i] = fw.iparams[i];
        wsh[1][i] = bw.iparams[i];
    }
    __syncthreads();
    const ModelViewT<R> &Sv = fwsearch ? fw : bw;
    bool act = false;
    u128 pat = 0;
    if (rl < nr) {
        const uint64_t r = r0 + rl;
        const uint32_t L = pq_read_len(G, r);
        act = D < L && Sv.nU > 0;
        if (act)
            pat = chain_pattern(fwt, pq_read_start(G, r), L, ch, D);
    }
    // every lane's prefix (zero where the chain takes no seed), the network for the
    // whole wave at once
    const uint64_t xk = (uint64_t) (pat >> (128 - 2 * d.k));
    const float y = mlp_forward_mfma(&wsh[fwsearch ? 0 : 1][0], d, xk);
    if (act)
        pred[(uint64_t) ch * nr + rl] = leaf_entry(y, Sv, leaves, S, xk);
}

// Block of 256 threads = 64 reads; wave q handles chain q of those reads, so the
// model of a wave is uniform.  Writes Seed{search interval} at depth D for
// chains of reads longer than D (size 0 when the pattern does not occur).
template <typename R>
// (launch bounds: 8 waves per SIMD asked for; the compiler settles at 96 vector
// registers, 5 waves per SIMD instead of 4, without spilling vector registers --
// more lanes in flight for this latency-bound kernel)
__global__ void __launch_bounds__(256, 8) k_learned_seed(Geom G, const uint64_t *fwt, uint64_t r0,
                                                      uint64_t nr, uint32_t D,
                                                      const uint64_t *pred,
                                                      ModelViewT<R> fw, ModelViewT<R> bw,
                                                      SeedT<R> *seeds, QStats *qs, ChainBuf cb,
                                                      uint32_t t, unsigned long long *stats)
{
    const int ch = (int) (threadIdx.x >> 6);
    const uint64_t rl = (uint64_t) blockIdx.x * 64 + (threadIdx.x & 63);
    const bool fwsearch = (ch == 0 || ch == 3);
    const int sdir = fwsearch ? 0 : 1;
    uint32_t probes[2] = {0, 0}, looks[2] = {0, 0};
    uint64_t win[2] = {0, 0};
    uint32_t absent = 0, verified = 0;
    uint32_t nrec = ~0u;   // records of a chain finished here
    // candidate staging; the block's seeds and record counts, written out in the
    // search kernels' order (read-major: 4 chains per read) in one coalesced pass
    __shared__ uint64_t stg[STAGE_SLOTS * STAGE_STRIDE];
    __shared__ SeedT<R> sds[256];
    __shared__ uint32_t nsh[256];
    const ModelViewT<R> &Sv = fwsearch ? fw : bw;
    bool act = false;
    u128 pat = 0;
    uint64_t g1 = 0;
    if (rl < nr) {
        const uint64_t r = r0 + rl;
        const uint32_t L = pq_read_len(G, r);
        act = D < L && Sv.nU > 0;
        if (act) {
            pat = chain_pattern(fwt, pq_read_start(G, r), L, ch, D);
            g1 = pred[(uint64_t) ch * nr + rl];   // the model's sampled row (k_learned_predict)
        }
    }
    if (rl < nr) {
        SeedT<R> sd = {0, SEED_NONE};
        if (act) {
            // the pattern's rows: the first row whose key is >= key, up to the first
            // row whose key is > key; the model predicts the sampled row and the
            // sampled keys bracket both ends to a stride of rows
            const u128 key = pat | 1u, above = key + 1u;   // bit 1 of a key is never set
            const uint64_t st = Sv.stride, ns = Sv.nU;
            const uint64_t jl = lower_bound_from(Sv.key, ns, key, g1, probes[sdir]);
            looks[sdir]++;
            win[sdir] = jl > g1 ? jl - g1 : g1 - jl;
            // sampled rows jl..jh-1 hold the pattern (none when jh == jl)
            const bool m1 = jl < ns && Sv.key[jl] == key;
            const uint64_t jh = m1 ? lower_bound_from(Sv.key, ns, above, jl, probes[sdir]) : jl;
            const uint32_t minsize = (ch == 0 || ch == 2) ? 2u : 1u;
            const uint64_t chain = rl * 4 + ch, r = r0 + rl;
            uint64_t lo, hi;
            bool done = false;
            if (cb.vmax > 0 && D <= t && (!m1 || (jh - jl - 1) * st + 1 <= cb.vmax)) {
                // the interval may be narrow enough to finish the chain on the text
                // (depth D <= t, nothing recorded yet): scan it outward from the
                // sampled rows, verifying its rows on the way
                const uint32_t L = pq_read_len(G, r);
                const ScanOut so = scan_chain<R>(
                    G, Sv.sa, Sv.N, fwsearch ? bw.t2 : fw.t2, ch < 2 ? bw.t2 : fw.t2,
                    (ch == 1 || ch == 3) ? ~0ull : 0ull, pq_read_start(G, r), L, D, key,
                    jl > 0 ? (jl - 1) * st + 1 : 0, jl < ns ? jl * st : Sv.N, m1, minsize, t,
                    cb.vmax, cb.lc + chain * cb.cap, cb.o + chain * cb.cap, stg + threadIdx.x);
                lo = so.lo;
                hi = so.hi;
                verified = so.visited;
                done = hi - lo >= minsize && hi - lo <= cb.vmax;
                if (done)
                    nrec = so.k;
            } else {
                // a wide interval: binary search on the suffix array within the
                // bracketing strides
                lo = row_lower_bound(G, Sv, jl > 0 ? (jl - 1) * st + 1 : 0,
                                     jl < ns ? jl * st : Sv.N, key, D, probes[sdir]);
                const uint64_t hlo = jh > 0 ? (jh - 1) * st + 1 : 0;
                hi = row_lower_bound(G, Sv, hlo > lo ? hlo : lo, jh < ns ? jh * st : Sv.N, above,
                                     D, probes[sdir]);
            }
            sd.l = (R) lo;
            // finished chains keep their records; an interval of 2^32 - 2 rows or
            // more takes the plain start instead
            sd.size = done                 ? SEED_DONE
                      : hi - lo < SEED_DONE ? (uint32_t) (hi - lo)
                                            : SEED_NONE;
            absent = hi == lo;   // no suffix starts with the pattern: the chain is empty
        }
        sds[(threadIdx.x & 63) * 4 + ch] = sd;
        nsh[(threadIdx.x & 63) * 4 + ch] = nrec;
    }
    if (stats) {
        uint64_t vr = verified;
        for (int o = 32; o > 0; o >>= 1)
            vr += __shfl_xor(vr, o, 64);
        if ((threadIdx.x & 63) == 0 && vr)
            wstat_add(stats, WS_SEED_ROWS, vr);
    }
    // block reduction of the statistics, one set of atomics per block
    __shared__ unsigned long long red[4][9];
    unsigned long long v[9] = {looks[0], looks[1], probes[0], probes[1], win[0],
                               win[1],   win[0],   win[1],    absent};
    for (int o = 32; o > 0; o >>= 1)
        for (int q = 0; q < 9; q++) {
            unsigned long long w = __shfl_xor(v[q], o, 64);
            v[q] = (q == 6 || q == 7) ? (w > v[q] ? w : v[q]) : v[q] + w;
        }
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 9; q++)
            red[ch][q] = v[q];
    __syncthreads();
    {
        // thread i of the block writes chain i of its 64 reads (rl * 4 + ch order);
        // chains the search kernel must still walk go to the work list
        const uint64_t i = (uint64_t) blockIdx.x * 256 + threadIdx.x;
        bool listed = false;
        if (i < nr * 4) {
            const SeedT<R> sd = sds[threadIdx.x];
            const uint32_t q = threadIdx.x & 3, ms = (q == 0 || q == 2) ? 2u : 1u;
            seeds[i] = sd;
            cb.vjob[i] = 0;
            listed = sd.size == SEED_NONE || (sd.size != SEED_DONE && sd.size >= ms);
            if (nsh[threadIdx.x] != ~0u)
                cb.n[i] = nsh[threadIdx.x];   // finished here
            else if (!listed)
                cb.n[i] = 0;                  // empty at depth D: no records
        }
        if (cb.wq) {
            // one atomic per block: the waves' counts, their prefix, one reservation
            __shared__ uint32_t wcnt[4], wbase;
            const uint64_t bal = __ballot(listed);
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            if (lane == 0)
                wcnt[wv] = (uint32_t) __popcll(bal);
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint32_t tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
                wbase = tot ? atomicAdd(&cb.wq_cnt[0], tot) : 0u;
            }
            __syncthreads();
            uint32_t base = wbase;
            for (int w = 0; w < wv; w++)
                base += wcnt[w];
            if (listed)
                cb.wq[base + (uint32_t) __popcll(bal & ((1ull << lane) - 1))] = (uint32_t) i;
        }
    }
    if (threadIdx.x < 9) {
        int q = threadIdx.x;
        bool mx = (q == 6 || q == 7);
        unsigned long long a = red[0][q];
        for (int w = 1; w < 4; w++)
            a = mx ? (red[w][q] > a ? red[w][q] : a) : a + red[w][q];
        // this block's copy of the statistics (spread as the work counters)
        QStats *qb = (QStats *) ((unsigned long long *) qs + (blockIdx.x % WSTAT_SLOTS) * WSTAT_STRIDE);
        unsigned long long *dst = q < 2   ? &qb->lookups[q]
                                  : q < 4 ? &qb->probes[q - 2]
                                  : q < 6 ? &qb->win_sum[q - 4]
                                  : q < 8 ? &qb->win_max[q - 6]
                                          : &qb->absent;
        if (a) {
            if (mx)
                atomicMax(dst, a);
            else
                atomicAdd(dst, a);
        }
    }
}


// ---- host driver -------------------------------------------------------------------------

void pq_learned_free(pq_ctx_t *ctx)
{
    if (!ctx->learned)
        return;
    for (int d = 0; d < 2; d++) {
        DirModel &M = ctx->learned->dir[d];
        pq_free(M.params);
        pq_free(M.iparams);
        pq_free(M.range);
        pq_free(M.knots);
    }
    pq_free(ctx->learned->qstats);
    delete ctx->learned;
    ctx->learned = nullptr;
}

extern "C" int pq_set_learned_keys(pq_ctx_t *ctx, uint32_t stride)
{
    if (!ctx || (stride & (stride - 1)) || stride > (1u << 16))
        return PQ_ERR_ARG;
    ctx->key_stride = stride;
    return PQ_OK;
}

extern "C" void pq_learned_default(pq_learned_cfg_t *cfg)
{
    if (!cfg)
        return;
    cfg->k = 21;
    cfg->hidden = 64;
    cfg->leaves = 4096;
    cfg->segments = 0;
    cfg->epochs = 1;
    cfg->batch = 16384;
    cfg->samples = 1u << 20;
    cfg->lr = 1e-5f;
    cfg->closed_form = 1;
    cfg->seed = 12345;
    cfg->sample_stride = 8;
}

static uint64_t host_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static float host_uniform(uint64_t seed, uint64_t i)
{
    // uniform in [-1, 1)
    return (float) (2.0 * (double) (host_mix(seed + 0x9E3779B97F4A7C15ull * (i + 1)) >> 11) *
                        0x1.0p-53 -
                    1.0);
}

// Solve (A + lambda I) w = b for symmetric positive semi-definite A (m x m, row
// major) by Cholesky in double; lambda = 1e-10 * max diagonal keeps it definite.
static bool solve_spd(std::vector<double> A, const std::vector<double> &b, uint32_t m,
                      std::vector<double> &w)
{
    double dmax = 0;
    for (uint32_t i = 0; i < m; i++)
        dmax = std::max(dmax, A[(size_t) i * m + i]);
    const double lam = 1e-10 * (dmax > 0 ? dmax : 1.0);
    for (uint32_t i = 0; i < m; i++)
        A[(size_t) i * m + i] += lam;
    for (uint32_t j = 0; j < m; j++) {
        double s = A[(size_t) j * m + j];
        for (uint32_t q = 0; q < j; q++)
            s -= A[(size_t) j * m + q] * A[(size_t) j * m + q];
        if (!(s > 0))
            return false;
        double djj = std::sqrt(s);
        A[(size_t) j * m + j] = djj;
        for (uint32_t i = j + 1; i < m; i++) {
            double t = A[(size_t) i * m + j];
            for (uint32_t q = 0; q < j; q++)
                t -= A[(size_t) i * m + q] * A[(size_t) j * m + q];
            A[(size_t) i * m + j] = t / djj;
        }
    }
    w.assign(m, 0.0);
    for (uint32_t i = 0; i < m; i++) {   // L y = b
        double t = b[i];
        for (uint32_t q = 0; q < i; q++)
            t -= A[(size_t) i * m + q] * w[q];
        w[i] = t / A[(size_t) i * m + i];
    }
    for (int i = (int) m - 1; i >= 0; i--) {   // L^T w = y
        double t = w[i];
        for (uint32_t q = i + 1; q < m; q++)
            t -= A[(size_t) q * m + i] * w[q];
        w[i] = t / A[(size_t) i * m + i];
    }
    return true;
}

// Mean squared error of the current parameters over the n samples.
static int sample_loss(pq_ctx_t *ctx, const float *params, MlpDims dims, const uint64_t *xs,
                       const float *targets, uint32_t n, float *part, float *lpart, double *loss)
{
    const uint32_t nb = (n + TRAIN_BLOCK - 1) / TRAIN_BLOCK;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TRAIN_BLOCK), 0, ctx->stream, params, dims, xs,
                       targets, n, 1.0f / n, part, lpart);
    std::vector<float> hl(nb);
    PQ_CHECK(hipGetLastError());
    PQ_CHECK(hipMemcpyAsync(hl.data(), lpart, (size_t) nb * 4, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    double s = 0;
    for (float v : hl)
        s += v;
    *loss = s / n;
    return PQ_OK;
}

// Inference copy of the parameters with W1 rows padded to MLP_K_MAX columns.
static int upload_inference_params(pq_ctx_t *ctx, const Learned &LM, DirModel &M)
{
    const uint32_t P = mlp_nparams(LM.k, LM.h);
    std::vector<float> p(P), q(MLP_P_MAX, 0.0f);
    PQ_CHECK(hipMemcpyAsync(p.data(), M.params, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < LM.h; i++)
        for (uint32_t j = 0; j < LM.k; j++)
            q[i * MLP_K_MAX + j] = p[i * LM.k + j];
    for (uint32_t i = 0; i < 2 * LM.h + 1; i++)
        q[LM.h * MLP_K_MAX + i] = p[LM.h * LM.k + i];
    PQ_CHECK(hipMemcpyAsync(M.iparams, q.data(), MLP_P_MAX * 4, hipMemcpyHostToDevice,
                            ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

// Sampled keys of direction d: one gather per stride rows into a scratch slot of
// the context that is kept between trainings (no allocation per training once
// it is large enough).
// The sampled keys of direction d: the ones the last build wrote when it was
// asked for this stride (pq_set_learned_keys), else gathered here from the
// suffix array and the text.
template <typename R> static int build_samples_t(pq_ctx_t *ctx, Learned &LM, int d)
{
    DirModel &M = LM.dir[d];
    DirIndex &Dx = ctx->dir[d];
    const uint64_t N = ctx->N, ns = (N + LM.stride - 1) / LM.stride;
    M.nU = ns;
    if (Dx.keys_stride == LM.stride && Dx.keys_D == LM.depth) {
        M.key = (u128 *) ctx->scratch[SLOT_SKEY0 + d].p;
        return PQ_OK;
    }
    Dx.keys_stride = Dx.keys_D = 0;
    M.key = (u128 *) ctx->tmp(SLOT_SKEY0 + d, (ns + 1) * sizeof(u128));
    if (!M.key)
        return ctx->fail(PQ_ERR_NOMEM, "learned index keys");
    if (ns)
        hipLaunchKernelGGL(k_sample_keys<R>, dim3(pq_blocks_gs(ns, 256)), dim3(256), 0,
                           ctx->stream, ctx->geom(), Dx.text2, (const R *) Dx.sa,
                           LM.stride, ns, LM.depth, M.key);
    PQ_CHECK(hipGetLastError());
    Dx.keys_stride = LM.stride;
    Dx.keys_D = LM.depth;
    return PQ_OK;
}

static int build_samples(pq_ctx_t *ctx, Learned &LM, int d)
{
    return ctx->wide ? build_samples_t<uint64_t>(ctx, LM, d) : build_samples_t<uint32_t>(ctx, LM, d);
}

static int train_dir(pq_ctx_t *ctx, Learned &LM, int d, const pq_learned_cfg_t &cfg)
{
    DirModel &M = LM.dir[d];
    const MlpDims dims = {LM.k, LM.h};
    const uint32_t P = mlp_nparams(LM.k, LM.h);
    const uint32_t S = cfg.samples;
    const uint32_t NL = LM.leaves, SG = LM.segs;
    int rc = build_samples(ctx, LM, d);
    if (rc)
        return rc;
    if ((rc = pq_alloc(ctx, (void **) &M.params, P * 4)) ||
        (rc = pq_alloc(ctx, (void **) &M.iparams, MLP_P_MAX * 4)) ||
        (rc = pq_alloc(ctx, (void **) &M.range, (size_t) NL * 16)) ||
        (rc = pq_alloc(ctx, (void **) &M.knots, (size_t) NL * (SG + 1) * 4)))
        return rc;
    // seeded initial parameters: W1 ~ U(-1/sqrt(k), 1/sqrt(k)); b1 = 5 for the
    // first half of the units (always active: affine features), 0 for the rest
    // (ReLU features); w2 ~ U(-1/sqrt(h), 1/sqrt(h)); b2 = 0.5
    std::vector<float> hp(P, 0.0f);
    const float s1 = 1.0f / std::sqrt((float) LM.k), s2 = 1.0f / std::sqrt((float) LM.h);
    const uint64_t ps = cfg.seed * 2 + d;
    for (uint32_t i = 0; i < LM.h * LM.k; i++)
        hp[i] = s1 * host_uniform(ps, i);
    for (uint32_t i = 0; i < LM.h; i++) {
        hp[LM.h * LM.k + i] = i < LM.h / 2 ? 5.0f : 0.0f;
        hp[LM.h * LM.k + LM.h + i] = s2 * host_uniform(ps, LM.h * LM.k + i);
    }
    hp[P - 1] = 0.5f;
    PQ_CHECK(hipMemcpyAsync(M.params, hp.data(), P * 4, hipMemcpyHostToDevice, ctx->stream));
    if (M.nU == 0) {
        if ((rc = upload_inference_params(ctx, LM, M)))
            return rc;
        // no suffix reaches the seed depth: every chain keeps the plain start
        PQ_CHECK(hipMemsetAsync(M.range, 0, (size_t) NL * 16, ctx->stream));
        PQ_CHECK(hipMemsetAsync(M.knots, 0, (size_t) NL * (SG + 1) * 4, ctx->stream));
        return PQ_OK;
    }
    const uint32_t nb = cfg.batch / TRAIN_BLOCK, nball = S / TRAIN_BLOCK;
    const uint32_t ngram = (S + GRAM_CHUNK - 1) / GRAM_CHUNK, m = LM.h + 2, ne = m * (m + 1) / 2;
    size_t need = (size_t) S * 8 + (size_t) S * 4 * 2 + (size_t) P * 4 * 2 +
                  (size_t) nball * P * 4 + (size_t) nball * 4 + (size_t) ngram * ne * 8 +
                  (size_t) ne * 8 + 4096;
    char *w = (char *) ctx->tmp(SLOT_LEARNED, need);
    if (!w)
        return ctx->fail(PQ_ERR_NOMEM, "learned index scratch");
    uint64_t *xs = (uint64_t *) w;
    float *targets = (float *) (xs + S);
    uint32_t *leaf = (uint32_t *) (targets + S);
    float *mom = (float *) (leaf + S);
    float *vel = mom + P;
    float *part = vel + P;
    float *lpart = part + (size_t) nball * P;
    double *gpart = (double *) (((uintptr_t) (lpart + nball) + 15) & ~(uintptr_t) 15);
    double *gram = gpart + (size_t) ngram * ne;
    hipLaunchKernelGGL(k_sample_rows, dim3(pq_blocks(S, 256)), dim3(256), 0, ctx->stream, M.key,
                       M.nU, LM.k, S, cfg.seed ^ (0xD1B54A32D192ED03ull * (d + 1)), xs, targets);
    PQ_CHECK(hipGetLastError());
    if (cfg.closed_form) {
        // output layer: least squares over the hidden activations of all samples
        hipLaunchKernelGGL(k_ls_gram, dim3(ngram), dim3(256), 0, ctx->stream, M.params, dims, xs,
                           targets, S, gpart);
        hipLaunchKernelGGL(k_sum_parts<double>, dim3(pq_blocks(ne, 256)), dim3(256), 0,
                           ctx->stream, gpart, ngram, ne, gram);
        PQ_CHECK(hipGetLastError());
        std::vector<double> hg(ne);
        PQ_CHECK(hipMemcpyAsync(hg.data(), gram, ne * 8, hipMemcpyDeviceToHost, ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
        const uint32_t q = LM.h + 1;   // unknowns: w2[0..h), b2
        auto g = [&](uint32_t i, uint32_t j) {
            if (i > j)
                std::swap(i, j);
            return hg[(size_t) i * m - (size_t) i * (i - 1) / 2 + (j - i)];   // row i of upper tri
        };
        std::vector<double> A((size_t) q * q), b(q), sol;
        for (uint32_t i = 0; i < q; i++) {
            for (uint32_t j = 0; j < q; j++)
                A[(size_t) i * q + j] = g(i, j);
            b[i] = g(i, m - 1);
        }
        if (solve_spd(A, b, q, sol)) {
            for (uint32_t i = 0; i < LM.h; i++)
                hp[LM.h * LM.k + LM.h + i] = (float) sol[i];
            hp[P - 1] = (float) sol[LM.h];
            PQ_CHECK(hipMemcpyAsync(M.params, hp.data(), P * 4, hipMemcpyHostToDevice,
                                    ctx->stream));
        }
    }
    if ((rc = sample_loss(ctx, M.params, dims, xs, targets, S, part, lpart, &M.fit_loss)))
        return rc;
    // minibatch Adam over the samples in a fixed order, step size decaying
    // linearly to 1% over the run
    PQ_CHECK(hipMemsetAsync(mom, 0, (size_t) P * 8, ctx->stream));
    const uint32_t steps = S / cfg.batch, total = steps * cfg.epochs;
    double b1t = 1.0, b2t = 1.0;
    for (uint32_t t = 0; t < total; t++) {
        uint32_t b = t % steps;
        b1t *= 0.9;
        b2t *= 0.999;
        float lr = (float) (cfg.lr * (1.0 - 0.99 * (double) t / (double) total));
        hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TRAIN_BLOCK), 0, ctx->stream, M.params, dims,
                           xs + (size_t) b * cfg.batch, targets + (size_t) b * cfg.batch,
                           cfg.batch, 1.0f / cfg.batch, part, lpart);
        hipLaunchKernelGGL(k_mlp_adam, dim3(pq_blocks(P, 256)), dim3(256), 0, ctx->stream,
                           M.params, mom, vel, part, nb, P, lr, (float) b1t, (float) b2t);
    }
    PQ_CHECK(hipGetLastError());
    if ((rc = sample_loss(ctx, M.params, dims, xs, targets, S, part, lpart, &M.loss)))
        return rc;
    if ((rc = upload_inference_params(ctx, LM, M)))
        return rc;
    // leaves: route every sample, x range per leaf
    hipLaunchKernelGGL(k_mlp_route, dim3(pq_blocks(S, 256)), dim3(256), 0, ctx->stream, M.params,
                       dims, xs, S, NL, leaf);
    PQ_CHECK(hipGetLastError());
    std::vector<uint64_t> hx(S);
    std::vector<uint32_t> hleaf(S);
    PQ_CHECK(hipMemcpyAsync(hx.data(), xs, (size_t) S * 8, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipMemcpyAsync(hleaf.data(), leaf, (size_t) S * 4, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<uint64_t> range(2 * (size_t) NL);
    std::vector<uint8_t> seen(NL, 0);
    for (uint32_t i = 0; i < S; i++) {
        uint32_t j = hleaf[i];
        if (!seen[j]) {
            seen[j] = 1;
            range[2 * j] = range[2 * j + 1] = hx[i];
        } else {
            range[2 * j] = std::min(range[2 * j], hx[i]);
            range[2 * j + 1] = std::max(range[2 * j + 1], hx[i]);
        }
    }
    // a leaf no sample reached keeps the single point where the previous one ends
    uint64_t prev = 0;
    for (uint32_t j = 0; j < NL; j++) {
        if (!seen[j])
            range[2 * j] = range[2 * j + 1] = prev;
        prev = range[2 * j + 1];
    }
    PQ_CHECK(hipMemcpyAsync(M.range, range.data(), range.size() * 8, hipMemcpyHostToDevice,
                            ctx->stream));
    hipLaunchKernelGGL(k_leaf_knots, dim3(NL), dim3(256), 0, ctx->stream, M.key, M.nU, LM.k,
                       M.range, SG, M.knots);
    PQ_CHECK(hipGetLastError());
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

template <typename R> static ModelViewT<R> model_view(const pq_ctx_t *ctx, const Learned &LM, int d)
{
    const DirModel &M = LM.dir[d];
    return {M.key,       M.nU,     LM.stride, ctx->N, (const R *) ctx->dir[d].sa,
            ctx->dir[d].text2, M.iparams, M.range, M.knots};
}

extern "C" int pq_learned_train(pq_ctx_t *ctx, const pq_learned_cfg_t *cfg_in)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (!ctx->dir[0].built || !ctx->dir[1].built)
        return ctx->fail(PQ_ERR_STATE, "index not built");
    pq_learned_cfg_t cfg;
    pq_learned_default(&cfg);
    if (cfg_in)
        cfg = *cfg_in;
    if (cfg.k == 0 || cfg.k > MLP_K_MAX || cfg.hidden == 0 || cfg.hidden > MLP_H_MAX ||
        cfg.leaves == 0 || cfg.leaves > (1u << 20) || cfg.segments > SEG_MAX || cfg.batch == 0 ||
        cfg.batch % TRAIN_BLOCK ||
        cfg.samples < cfg.batch || cfg.samples % TRAIN_BLOCK || cfg.sample_stride == 0 ||
        cfg.sample_stride > 4096)
        return ctx->fail(PQ_ERR_ARG, "learned index configuration out of range");
    if (ctx->N == 0 || ctx->t < 2)
        return ctx->fail(PQ_ERR_STATE, "no suffixes to learn (empty text or t < 2)");
    PQ_CHECK(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    pq_learned_free(ctx);
    Learned *LM = new Learned();
    // seed depth: up to 63 codes (one 128-bit key) but never past the minimum
    // overlap, since candidates are recorded from depth t on; the model sees the
    // first k <= D codes
    LM->depth = std::min<uint32_t>(63, ctx->t);
    LM->k = std::min<uint32_t>(cfg.k, LM->depth);
    LM->h = cfg.hidden;
    LM->leaves = cfg.leaves;
    // automatic piece count: about SEG_AUTO_ROWS sampled rows per piece, so that
    // interpolation inside a piece is off by a few sampled rows at most
    const uint64_t ns = (ctx->N + cfg.sample_stride - 1) / cfg.sample_stride;
    const uint64_t sauto = (ns + (uint64_t) cfg.leaves * SEG_AUTO_ROWS - 1) /
                           ((uint64_t) cfg.leaves * SEG_AUTO_ROWS);
    LM->segs = cfg.segments ? cfg.segments
                            : (uint32_t) std::min<uint64_t>(SEG_MAX, std::max<uint64_t>(16, sauto));
    LM->stride = (uint32_t) cfg.sample_stride;
    ctx->learned = LM;
    ctx->learned_gen++;
    int rc = pq_alloc(ctx, (void **) &LM->qstats, QSTATS_BYTES);
    for (int d = 0; d < 2 && rc == PQ_OK; d++)
        rc = train_dir(ctx, *LM, d, cfg);
    if (rc == PQ_OK && hipMemsetAsync(LM->qstats, 0, QSTATS_BYTES, ctx->stream) != hipSuccess)
        rc = ctx->fail(PQ_ERR_HIP, "clearing learned statistics");
    if (rc == PQ_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = ctx->fail(PQ_ERR_HIP, "learned index training");
    if (rc) {
        pq_learned_free(ctx);
        return rc;
    }
    LM->train_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PQ_OK;
}

extern "C" int pq_learned_stats(pq_ctx_t *ctx, pq_learned_stats_t *st)
{
    if (!ctx || !st)
        return PQ_ERR_ARG;
    Learned *LM = ctx->learned;
    if (!LM)
        return ctx->fail(PQ_ERR_STATE, "learned index not trained");
    // the copies summed (maxima for the windows)
    std::vector<unsigned long long> all(QSTATS_BYTES / 8);
    PQ_CHECK(hipMemcpyAsync(all.data(), LM->qstats, QSTATS_BYTES, hipMemcpyDeviceToHost,
                            ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    QStats q;
    memset(&q, 0, sizeof(q));
    for (uint32_t b = 0; b < WSTAT_SLOTS; b++) {
        QStats c;
        memcpy(&c, all.data() + (size_t) b * WSTAT_STRIDE, sizeof(c));
        for (int d = 0; d < 2; d++) {
            q.lookups[d] += c.lookups[d];
            q.probes[d] += c.probes[d];
            q.win_sum[d] += c.win_sum[d];
            q.win_max[d] = std::max(q.win_max[d], c.win_max[d]);
        }
        q.absent += c.absent;
    }
    memset(st, 0, sizeof(*st));
    for (int d = 0; d < 2; d++) {
        st->lookups[d] = q.lookups[d];
        st->probes[d] = q.probes[d];
        st->mean_window[d] = q.lookups[d] ? (double) q.win_sum[d] / (double) q.lookups[d] : 0.0;
        st->max_window[d] = q.win_max[d];
        st->train_loss[d] = LM->dir[d].loss;
        st->fit_loss[d] = LM->dir[d].fit_loss;
        st->table_size[d] = LM->dir[d].nU;
    }
    st->absent = q.absent;
    st->fallback = 0;   // every seed comes from its own table: none falls back
    st->train_ms = LM->train_ms;
    st->depth = LM->depth;
    st->k = LM->k;
    st->segments = LM->segs;
    return PQ_OK;
}

// Launch the seed kernel for reads [r0, r0 + nr) of a search batch (stream
// ordered, no synchronisation).  Statistics accumulate over one search.
template <typename R>
static void launch_seeds(pq_ctx_t *ctx, const Learned *LM, uint64_t r0, uint64_t nr,
                         SeedT<R> *d_seeds, const ChainBuf &cb, unsigned long long *stats,
                         hipEvent_t mid)
{
    ModelViewT<R> v[2] = {model_view<R>(ctx, *LM, 0), model_view<R>(ctx, *LM, 1)};
    MlpDims dims = {LM->k, LM->h};
    // the predictions go to the chains' verification-start words (cb.vl, 8 bytes
    // per chain), which the search kernels only write after the seed kernel
    hipLaunchKernelGGL(k_learned_predict<R>, dim3(pq_blocks(nr, 64)), dim3(256), 0, ctx->stream,
                       ctx->geom(), ctx->dir[0].text2, r0, nr, LM->depth, dims, LM->leaves,
                       LM->segs, v[0], v[1], cb.vl);
    if (mid)
        (void) hipEventRecord(mid, ctx->stream);
    // the seed kernel finishes chains by its own threshold (the search kernel's is
    // cb.vmax)
    ChainBuf cbs = cb;
    cbs.vmax = std::min(std::min<uint32_t>(ctx->seed_rows, VERIFY_MAX), cb.cap);
    hipLaunchKernelGGL(k_learned_seed<R>, dim3(pq_blocks(nr, 64)), dim3(256), 0, ctx->stream,
                       ctx->geom(), ctx->dir[0].text2, r0, nr, LM->depth, cb.vl, v[0], v[1],
                       d_seeds, LM->qstats, cbs, ctx->t, stats);
}

int pq_learned_seeds(pq_ctx_t *ctx, uint64_t r0, uint64_t nr, void *d_seeds, uint32_t *kseed,
                     const ChainBuf &cb, unsigned long long *stats, hipEvent_t mid)
{
    Learned *LM = ctx->learned;
    if (!LM)
        return ctx->fail(PQ_ERR_STATE, "learned index not trained");
    if (r0 == 0)   // a kernel, not a memset, so that the search can be captured
        hipLaunchKernelGGL(k_wstat_clear, dim3(WSTAT_SLOTS), dim3(64), 0, ctx->stream,
                           (unsigned long long *) LM->qstats,
                           (uint32_t) (sizeof(QStats) / sizeof(unsigned long long)));
    if (ctx->wide)
        launch_seeds<uint64_t>(ctx, LM, r0, nr, (SeedT<uint64_t> *) d_seeds, cb, stats, mid);
    else
        launch_seeds<uint32_t>(ctx, LM, r0, nr, (SeedT<uint32_t> *) d_seeds, cb, stats, mid);
    PQ_CHECK(hipGetLastError());
    *kseed = LM->depth;
    return PQ_OK;
}

// Gradient of (1/n) sum (MLP(x) - y)^2 and the loss, by the training kernels, for
// n host samples given as k-code prefix values (checked against the CPU oracle).
extern "C" int pq_mlp_gradients(pq_ctx_t *ctx, uint32_t k, uint32_t hidden, const float *params,
                                const uint64_t *xs, const float *targets, uint32_t n, float *grad,
                                float *loss)
{
    if (!ctx || !params || !xs || !targets || !grad || n == 0 || k == 0 || k > MLP_K_MAX ||
        hidden == 0 || hidden > MLP_H_MAX)
        return PQ_ERR_ARG;
    PQ_CHECK(hipSetDevice(ctx->device));
    const uint32_t P = mlp_nparams(k, hidden);
    const uint32_t nb = (n + TRAIN_BLOCK - 1) / TRAIN_BLOCK;
    size_t need = (size_t) n * 8 + (size_t) P * 4 * 2 + (size_t) n * 4 + (size_t) nb * P * 4 +
                  (size_t) nb * 4 + 1024;
    char *w = (char *) ctx->tmp(SLOT_LEARNED + 1, need);
    if (!w)
        return ctx->fail(PQ_ERR_NOMEM, "gradient scratch");
    uint64_t *dx = (uint64_t *) w;
    float *dp = (float *) (dx + n), *dg = dp + P, *dt = dg + P;
    float *part = dt + n, *lpart = part + (size_t) nb * P;
    PQ_CHECK(hipMemcpyAsync(dp, params, P * 4, hipMemcpyHostToDevice, ctx->stream));
    PQ_CHECK(hipMemcpyAsync(dx, xs, (size_t) n * 8, hipMemcpyHostToDevice, ctx->stream));
    PQ_CHECK(hipMemcpyAsync(dt, targets, (size_t) n * 4, hipMemcpyHostToDevice, ctx->stream));
    MlpDims dims = {k, hidden};
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TRAIN_BLOCK), 0, ctx->stream, dp, dims, dx, dt, n,
                       1.0f / n, part, lpart);
    hipLaunchKernelGGL(k_sum_parts<float>, dim3(pq_blocks(P, 256)), dim3(256), 0, ctx->stream,
                       part, nb, P, dg);
    PQ_CHECK(hipGetLastError());
    std::vector<float> lp(nb);
    PQ_CHECK(hipMemcpyAsync(grad, dg, P * 4, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipMemcpyAsync(lp.data(), lpart, nb * 4, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    if (loss) {
        double s = 0;
        for (float v : lp)
            s += v;
        *loss = (float) (s / n);
    }
    return PQ_OK;
}

// The trained network of one direction on n prefix values: the search's matrix-
// core form (mfma = 1: one wave per 64 values, inference parameters in LDS) or
// the scalar padded form (mfma = 0).
__global__ void __launch_bounds__(64) k_mlp_eval(const float *iparams, MlpDims d,
                                                 const uint64_t *xs, uint32_t n, int mfma,
                                                 float *out)
{
    __shared__ float sp[MLP_P_MAX];
    for (uint32_t i = threadIdx.x; i < MLP_P_MAX; i += blockDim.x)
        sp[i] = iparams[i];
    __syncthreads();
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const uint64_t x = i < n ? xs[i] : 0;
    float y;
    if (mfma)
        y = mlp_forward_mfma(sp, d, x);          // every lane of the wave takes part
    else
        y = mlp_forward_padded(sp, d, x);
    if (i < n)
        out[i] = y;
}

extern "C" int pq_mlp_eval(pq_ctx_t *ctx, int dir, const uint64_t *xs, uint32_t n, int mfma,
                           float *out)
{
    if (!ctx || dir < 0 || dir > 1 || (n && (!xs || !out)))
        return PQ_ERR_ARG;
    Learned *LM = ctx->learned;
    if (!LM)
        return ctx->fail(PQ_ERR_STATE, "learned index not trained");
    if (n == 0)
        return PQ_OK;
    PQ_CHECK(hipSetDevice(ctx->device));
    char *w = (char *) ctx->tmp(SLOT_LEARNED + 1, (size_t) n * 12 + 256);
    if (!w)
        return ctx->fail(PQ_ERR_NOMEM, "evaluation scratch");
    uint64_t *dx = (uint64_t *) w;
    float *dy = (float *) (dx + n);
    PQ_CHECK(hipMemcpyAsync(dx, xs, (size_t) n * 8, hipMemcpyHostToDevice, ctx->stream));
    MlpDims dims = {LM->k, LM->h};
    hipLaunchKernelGGL(k_mlp_eval, dim3((n + 63) / 64), dim3(64), 0, ctx->stream,
                       LM->dir[dir].iparams, dims, dx, n, mfma, dy);
    PQ_CHECK(hipGetLastError());
    PQ_CHECK(hipMemcpyAsync(out, dy, (size_t) n * 4, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

// Parameters of the trained model of direction dir (0 fw, 1 bw), host copy.
extern "C" int pq_learned_params(pq_ctx_t *ctx, int dir, float *params, uint32_t *count)
{
    if (!ctx || !count || dir < 0 || dir > 1)
        return PQ_ERR_ARG;
    Learned *LM = ctx->learned;
    if (!LM)
        return ctx->fail(PQ_ERR_STATE, "learned index not trained");
    uint32_t P = mlp_nparams(LM->k, LM->h);
    if (params) {
        if (*count < P)
            return ctx->fail(PQ_ERR_ARG, "parameter buffer too small");
        PQ_CHECK(hipMemcpyAsync(params, LM->dir[dir].params, P * 4, hipMemcpyDeviceToHost,
                                ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
    }
    *count = P;
    return PQ_OK;
}

// Host copy of the sampled keys of direction dir: per sample the key as two
// 64-bit halves (high first; first code in the top bits, bit 0 of the low half
// set when the suffix has at least D codes), its suffix-array row (j * stride)
// and the rows it stands for (the stride; fewer for the last).  *count in:
// capacity, out: samples (any output may be NULL to query it).
extern "C" int pq_learned_table(pq_ctx_t *ctx, int dir, uint64_t *keys, uint64_t *first,
                                uint32_t *rows, uint64_t *count)
{
    if (!ctx || !count || dir < 0 || dir > 1)
        return PQ_ERR_ARG;
    Learned *LM = ctx->learned;
    if (!LM)
        return ctx->fail(PQ_ERR_STATE, "learned index not trained");
    const DirModel &M = LM->dir[dir];
    if ((keys || first || rows) && *count < M.nU)
        return ctx->fail(PQ_ERR_ARG, "table buffer too small");
    *count = M.nU;
    if (M.nU == 0)
        return PQ_OK;
    if (keys) {
        std::vector<u128> k(M.nU);
        PQ_CHECK(hipMemcpyAsync(k.data(), M.key, M.nU * sizeof(u128), hipMemcpyDeviceToHost,
                                ctx->stream));
        PQ_CHECK(hipStreamSynchronize(ctx->stream));
        for (uint64_t i = 0; i < M.nU; i++) {
            keys[2 * i] = (uint64_t) (k[i] >> 64);
            keys[2 * i + 1] = (uint64_t) k[i];
        }
    }
    for (uint64_t i = 0; i < M.nU; i++) {
        const uint64_t r = i * LM->stride;
        if (first)
            first[i] = r;
        if (rows)
            rows[i] = (uint32_t) std::min<uint64_t>(LM->stride, ctx->N - r);
    }
    return PQ_OK;
}
