// pq_dev.h -- device-side data layout shared by the index build and the overlap
// search kernels (gfx950, wave64).
//
// HBM layout of one direction of the index (forward text = the reads; backward
// text = every read reversed, not complemented, as in sufsort.c:80-121):
//
//   text2  : 2-bit packed text, 32 codes per 64-bit word, code j of word w in
//            bits [63-2j-1, 63-2j] (most significant first) so that a word
//            compares like the string it holds.  A=0 C=1 G=2 T=3.  One spare
//            word at the end so that any 32-code window can be read with two
//            loads.
//   ends   : 1 bit per text position, set on the last code of every read
//            (shared by both directions: a reversed read occupies the same
//            range as the read).
//   sa     : suffix array: positions of row type R, uint32_t while the text
//            has fewer than 2^32 codes, uint64_t beyond (rows of the BWT and
//            occ tables take the same type).
//   occ    : one 64-byte OccLine per 128 BWT rows -- three bit planes of the
//            3-bit row symbol ($=0 A=1 C=2 G=3 T=4, padding 7) plus the counts
//            of A, C, G, T before the line.  A backward-search step touches the
//            lines of l and u+1 only (one line when both fall in it), and a
//            line holds every symbol, so a bidirectional step needs no other
//            memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PQ_WAVE 64

struct __attribute__((aligned(64))) OccLine {
    uint64_t p0[2], p1[2], p2[2];
    uint32_t cnt[4];
};
static_assert(sizeof(OccLine) == 64, "OccLine must be one 64-byte line");

// Everything a search kernel needs about one direction, for row type R (the
// search reads the packed text, the occurrence lines and reads_index; not the
// suffix array).
template <typename R> struct DirViewT {
    const uint64_t *text2;
    const OccLine *occ;
    const uint32_t *reads_index;
    const R *sa;
    R cf[5];                 // backward-step base: C[c] + fix[c]; cf[0] = 0
};

// A row with a count (an interval start and its size, a table entry's first
// row and run length): 8 bytes for 32-bit rows, 16 for 64-bit rows, one load.
template <typename R> struct __attribute__((aligned(2 * sizeof(R)))) RowCount {
    R x;
    uint32_t y;
};

// C[0..5] of index.c:366-372 in the row type.
template <typename R> struct CTab {
    R c[6];
};

// Element of C selected by a per-lane symbol without indexing (which would
// move the table to LDS or scratch).
template <typename R> __device__ __forceinline__ R pq_pick6(const R (&v)[6], unsigned c)
{
    R x = v[0];
#pragma unroll
    for (unsigned q = 1; q < 6; q++)
        x = c == q ? v[q] : x;
    return x;
}

template <typename R> __device__ __forceinline__ R pq_pick5(const R (&v)[5], unsigned c)
{
    R x = v[0];
#pragma unroll
    for (unsigned q = 1; q < 5; q++)
        x = c == q ? v[q] : x;
    return x;
}

// Grid-stride loop bounds for kernels over 2^32 elements or more: a launch
// covers at most 2^32 work-items per dimension.
#define PQ_GRID_STRIDE_FOR(i, count)                                                      \
    for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < (count);   \
         i += (uint64_t) gridDim.x * blockDim.x)

// Work counters, spread over WSTAT_SLOTS copies WSTAT_STRIDE words apart (4 KiB
// + 128 B: different lines and memory channels).  A device-scope atomic is
// resolved on the memory side of this GPU (each XCD has its own L2), where
// atomics to one address serialize at roughly 10 ns each: a million of them
// from one launch to one counter cost about 10 ms.  Block b adds to copy
// b % WSTAT_SLOTS; the copies are summed once at the end.
#define WSTAT_SLOTS 64
#define WSTAT_STRIDE 528
#define WSTAT_WORDS 16           // counters per copy
enum {
    WS_POSITIONS = 0,            // rank positions of backward steps
    WS_LINES,                    // occurrence lines loaded by them
    WS_VERIFY_ROWS,              // rows finished on the text by k_verify
    WS_SEED_ROWS,                // rows the learned seed kernel read on the text
    WS_ITEMS,                    // candidate items expanded by the extraction
    WS_TIER_READS,               // + tier: reads finished by extraction tier 0..3
};

__device__ __forceinline__ void wstat_add(unsigned long long *ws, int q, unsigned long long v)
{
    if (v)
        atomicAdd(ws + (blockIdx.x % WSTAT_SLOTS) * WSTAT_STRIDE + q, v);
}

// Zeroes counters [0, words) of every copy (grid: WSTAT_SLOTS blocks of >= words).
static __global__ void k_wstat_clear(unsigned long long *ws, uint32_t words)
{
    if (threadIdx.x < words)
        ws[(uint64_t) blockIdx.x * WSTAT_STRIDE + threadIdx.x] = 0;
}

// out[q] = counter q summed over the copies, q < words (one block of >= words).
static __global__ void k_wstat_sum(const unsigned long long *ws, uint32_t words,
                                   unsigned long long *out)
{
    const uint32_t q = threadIdx.x;
    if (q >= words)
        return;
    unsigned long long v = 0;
    for (uint32_t b = 0; b < WSTAT_SLOTS; b++)
        v += ws[(uint64_t) b * WSTAT_STRIDE + q];
    out[q] = v;
}

// The value of lane l ^ J.  J = 1, 2: a DPP quad permutation inside the VALU;
// J < 32: ds_swizzle in bit-mask mode (xor within 32 lanes, no address
// arithmetic); 32: ds_bpermute.
template <int J> __device__ __forceinline__ uint32_t xor_lane(uint32_t v)
{
    if constexpr (J == 1)
        return (uint32_t) __builtin_amdgcn_mov_dpp((int) v, 0xB1, 0xF, 0xF, true);
    else if constexpr (J == 2)
        return (uint32_t) __builtin_amdgcn_mov_dpp((int) v, 0x4E, 0xF, 0xF, true);
    else if constexpr (J < 32)
        return (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, (J << 10) | 0x1F);
    else
        return (uint32_t) __shfl_xor((int) v, J, 64);
}

template <int J> __device__ __forceinline__ uint64_t xor_lane64(uint64_t v)
{
    return ((uint64_t) xor_lane<J>((uint32_t) (v >> 32)) << 32) | xor_lane<J>((uint32_t) v);
}

typedef unsigned __int128 u128;

// Read geometry: either a fixed read length or per-read offsets + end bitmap.
struct Geom {
    uint64_t n;              // reads
    uint64_t N;              // text length
    uint32_t L;              // fixed length, or 0 when lengths vary
    uint32_t Lmax;
    const uint64_t *off;     // n+1 read offsets
    const uint64_t *ends;    // end-of-read bitmap
    const uint32_t *end_rank;// #reads ending before each 64-position word
};

__device__ __forceinline__ uint64_t pq_window(const uint64_t *t2, uint64_t g)
{
    // 32 codes starting at text position g, most significant first.  Both words
    // are always loaded and joined by a branch-free funnel shift (the packed
    // text carries slack words past its end): a variant that loaded the second
    // word only for unaligned g, under an exec mask into registers the first
    // word had just been moved into, returned wrong windows now and then on
    // gfx950.
    uint64_t w = g >> 5;
    unsigned s = (unsigned) (g & 31) * 2;
    return (t2[w] << s) | ((t2[w + 1] >> 1) >> (63 - s));
}

__device__ __forceinline__ unsigned pq_code(const uint64_t *t2, uint64_t g)
{
    return (unsigned) (t2[g >> 5] >> (62 - 2 * (g & 31))) & 3u;
}

// Read containing text position g.  Fixed lengths: g / L by a double-precision
// reciprocal (g < 2^53), corrected to the exact quotient (a 64-bit division is a
// long instruction sequence on the device).
__device__ __forceinline__ uint64_t pq_read_of(const Geom &G, uint64_t g)
{
    if (G.L) {
        uint64_t q = (uint64_t) ((double) g * (1.0 / (double) G.L));
        const uint64_t L = G.L;
        if (q * L > g)
            q--;
        else if ((q + 1) * L <= g)
            q++;
        return q;
    }
    uint64_t w = g >> 6;
    uint64_t below = G.ends[w] & ((1ull << (g & 63)) - 1ull);
    return G.end_rank[w] + (uint64_t) __popcll(below);
}

__device__ __forceinline__ uint64_t pq_read_start(const Geom &G, uint64_t r)
{
    return G.L ? r * (uint64_t) G.L : G.off[r];
}

__device__ __forceinline__ uint32_t pq_read_len(const Geom &G, uint64_t r)
{
    return G.L ? G.L : (uint32_t) (G.off[r + 1] - G.off[r]);
}

// Codes remaining from g to the end of its read (>= 1).
__device__ __forceinline__ uint64_t pq_rem(const Geom &G, uint64_t g)
{
    if (G.L)
        return G.L - (uint32_t) (g - pq_read_of(G, g) * G.L);
    uint64_t r = pq_read_of(G, g);
    return G.off[r + 1] - g;
}

__device__ __forceinline__ bool pq_is_start(const Geom &G, uint64_t g)
{
    if (G.L)
        return pq_read_of(G, g) * G.L == g;   // no 64-bit division
    if (g == 0)
        return true;
    uint64_t p = g - 1;
    return (G.ends[p >> 6] >> (p & 63)) & 1ull;
}

// ---- occurrence lines -------------------------------------------------------

struct Ranks {
    uint32_t r[5];           // rank of $,A,C,G,T before the position
};

__device__ __forceinline__ void pq_line_ranks(const OccLine &ln, uint64_t line_idx,
                                              unsigned j, Ranks &out)
{
    uint64_t m0 = j >= 64 ? ~0ull : ((1ull << j) - 1ull);
    uint64_t m1 = j > 64 ? ((1ull << (j - 64)) - 1ull) : 0ull;
    uint64_t a0 = ln.p0[0], a1 = ln.p0[1];
    uint64_t b0 = ln.p1[0], b1 = ln.p1[1];
    uint64_t c0 = ln.p2[0], c1 = ln.p2[1];
    uint32_t nA = __popcll(a0 & ~b0 & ~c0 & m0) + __popcll(a1 & ~b1 & ~c1 & m1);
    uint32_t nC = __popcll(~a0 & b0 & ~c0 & m0) + __popcll(~a1 & b1 & ~c1 & m1);
    uint32_t nG = __popcll(a0 & b0 & ~c0 & m0) + __popcll(a1 & b1 & ~c1 & m1);
    uint32_t nT = __popcll(~a0 & ~b0 & c0 & m0) + __popcll(~a1 & ~b1 & c1 & m1);
    uint32_t nD = __popcll(~a0 & ~b0 & ~c0 & m0) + __popcll(~a1 & ~b1 & ~c1 & m1);
    uint32_t base = (uint32_t) (line_idx * 128u);
    uint32_t sum = ln.cnt[0] + ln.cnt[1] + ln.cnt[2] + ln.cnt[3];
    out.r[0] = base - sum + nD;
    out.r[1] = ln.cnt[0] + nA;
    out.r[2] = ln.cnt[1] + nC;
    out.r[3] = ln.cnt[2] + nG;
    out.r[4] = ln.cnt[3] + nT;
}

// What one backward step by symbol c (1..4) needs from a line, before row j of
// it: rows holding c and '$' rows.  Masks by selects on c, so nothing is
// indexed per lane.
struct StepRanks {
    uint32_t c, d;
};

__device__ __forceinline__ StepRanks pq_step_ranks(const OccLine &ln, uint64_t line_idx,
                                                   unsigned j, unsigned c)
{
    uint64_t m0 = j >= 64 ? ~0ull : ((1ull << j) - 1ull);
    uint64_t m1 = j > 64 ? ((1ull << (j - 64)) - 1ull) : 0ull;
    const uint64_t a0 = ln.p0[0], a1 = ln.p0[1];
    const uint64_t b0 = ln.p1[0], b1 = ln.p1[1];
    const uint64_t c0 = ln.p2[0], c1 = ln.p2[1];
    // symbol c: plane bits (c & 1, c & 2, c & 4)
    const bool s0 = c & 1u, s1 = c & 2u, s2 = c & 4u;
    uint64_t e0 = (s0 ? a0 : ~a0) & (s1 ? b0 : ~b0) & (s2 ? c0 : ~c0);
    uint64_t e1 = (s0 ? a1 : ~a1) & (s1 ? b1 : ~b1) & (s2 ? c1 : ~c1);
    uint64_t d0 = ~(a0 | b0 | c0), d1 = ~(a1 | b1 | c1);
    const uint32_t k0 = ln.cnt[0], k1 = ln.cnt[1], k2 = ln.cnt[2], k3 = ln.cnt[3];
    // '$' rows before the line: rows before it minus A..T rows before it
    const uint32_t dbase = (uint32_t) (line_idx * 128u) - (k0 + k1 + k2 + k3);
    const uint32_t kc = c == 1 ? k0 : c == 2 ? k1 : c == 3 ? k2 : k3;
    StepRanks r;
    r.c = kc + __popcll(e0 & m0) + __popcll(e1 & m1);
    r.d = dbase + __popcll(d0 & m0) + __popcll(d1 & m1);
    return r;
}

__device__ __forceinline__ void pq_load_line(const OccLine *occ, uint64_t idx, OccLine &ln)
{
    const uint4 *p = reinterpret_cast<const uint4 *>(occ + idx);
    uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
    ln.p0[0] = ((uint64_t) q0.y << 32) | q0.x;
    ln.p0[1] = ((uint64_t) q0.w << 32) | q0.z;
    ln.p1[0] = ((uint64_t) q1.y << 32) | q1.x;
    ln.p1[1] = ((uint64_t) q1.w << 32) | q1.z;
    ln.p2[0] = ((uint64_t) q2.y << 32) | q2.x;
    ln.p2[1] = ((uint64_t) q2.w << 32) | q2.z;
    ln.cnt[0] = q3.x;
    ln.cnt[1] = q3.y;
    ln.cnt[2] = q3.z;
    ln.cnt[3] = q3.w;
}

// Ranks of all symbols before pos.
__device__ __forceinline__ void pq_ranks(const OccLine *occ, uint64_t pos, Ranks &out)
{
    OccLine ln;
    pq_load_line(occ, pos >> 7, ln);
    pq_line_ranks(ln, pos >> 7, (unsigned) (pos & 127), out);
}

// Ranks before l and before u1 (= u+1); one line load when both share a line.
__device__ __forceinline__ void pq_ranks2(const OccLine *occ, uint64_t l, uint64_t u1,
                                          Ranks &rl, Ranks &ru)
{
    OccLine ln;
    uint64_t wl = l >> 7, wu = u1 >> 7;
    pq_load_line(occ, wl, ln);
    pq_line_ranks(ln, wl, (unsigned) (l & 127), rl);
    if (wu != wl)
        pq_load_line(occ, wu, ln);
    pq_line_ranks(ln, wu, (unsigned) (u1 & 127), ru);
}

// ---- keys of the learned index -------------------------------------------------------

// 64 codes from text position g, first code in the top bits.
__device__ __forceinline__ u128 text_codes(const uint64_t *t2, uint64_t g)
{
    return ((u128) pq_window(t2, g) << 64) | (u128) pq_window(t2, g + 32);
}

// Key bits kept for D codes: the top 2D bits.
__device__ __forceinline__ u128 top_mask(uint32_t D)
{
    return ~(u128) 0 << (128 - 2 * D);
}

// Key of the suffix at text position g: its first D codes (codes past the end of
// its read read as 0) | (the suffix has at least D codes).  Non-decreasing along
// the suffix array: a suffix shorter than D codes that is a prefix of a pattern
// sorts before the pattern's rows, as in the suffix array.
__device__ __forceinline__ u128 suffix_key(const Geom &G, const uint64_t *t2, uint64_t g,
                                           uint32_t D)
{
    const uint64_t rem = pq_rem(G, g);
    // up to 32 codes come from one 64-bit window (two words instead of three)
    u128 w = D <= 32 ? (u128) pq_window(t2, g) << 64 : text_codes(t2, g);
    w &= top_mask(D);
    if (rem < D)
        w &= ~(~(u128) 0 >> (2 * rem));
    return w | (u128) (rem >= D ? 1u : 0u);
}

