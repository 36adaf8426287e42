// pq_search.hip -- FM-index overlap search (string_graph.c:162-731) on gfx950.
//
// Candidates (find_intervals, string_graph.c:162-383).  Four independent chains
// per read: two per lane (the two of one index) on two adjacent lanes (k_find2),
// or one per lane on four adjacent lanes, each lane loading its own lines
// (k_find1) or the four lanes loading each line together (k_find1q):
//   chain 0 FB  pattern n[L-o..L)      searched in the forward index
//   chain 1 FF  comp(n[L-o..L))        searched in the backward index
//   chain 2 BF  rev(n[0..o))           searched in the backward index
//   chain 3 BB  rc(n[0..o))            searched in the forward index
// Each chain is a backward search from one code (or from the learned seed);
// every step costs the OccLines of l and u+1 only.  For t <= o < L the '$' rows
// of the interval are the reads whose text in the searched index begins with
// the pattern (string_graph.c:220-239 and the like take them by the same '$'
// step), and the chain records (o, '$' rank of the interval, count): reads
// reads_index[rank .. rank + count) of the searched index.  FB and BF stop when
// only the read itself is left (size < 2, string_graph.c:226,311), FF and BB
// when the interval empties (269,354).
//
// Irreducible edges (extract_{fw,bw}_edges, string_graph.c:386-731), restated
// on explicit strings.  The candidates of one side are expanded into items
// (type, o, m) through reads_index; each item's extension key is
// m[o..) for FB/BB and rc(m)[o..) for FF/BF (2-bit codes, so comparing words
// is comparing strings).  The reference's depth-first walk emits, in key order,
// exactly the keys that have no other key as a proper prefix; at such a key it
// takes the first candidate group (type order FB<FF, BF<BB, then o) and emits
// its reads other than n in read order (reads_index order of identical
// strings).  A group of lanes per read sorts the items with their keys in LDS
// and reads these rules off the sorted order (extract_sorted).
#include "pq_ctx.h"
#include "pq_chain.h"

#include <chrono>
#include <cstring>

// extraction tiers by items per read side: most reads fit the first (small LDS
// footprint, high occupancy); the few larger ones go to one wave per block with a
// large LDS share, and the rest to global-memory scratch
#define EXT_WAVES 4
// first-tier shapes the autotuner chooses from: lanes per read, items per side
static const uint32_t EXT_SHAPES[3][2] = {{4, 16}, {8, 32}, {16, 64}};
#define EXT_NSHAPES 3
#define KMAX_LDS2 512
#define KMAX_GLOBAL 65536
#define GLOBAL_SLOTS 256

// Orders the LDS accesses of a wave: what one lane wrote before it is visible
// to every lane after it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned chain_sym(const uint64_t *fwt, uint64_t s, uint32_t L, int ch,
                                              uint32_t o)
{
    bool from_end = ch < 2;
    uint64_t idx = from_end ? (uint64_t) (L - o) : (uint64_t) (o - 1);
    unsigned c = pq_code(fwt, s + idx);
    if (ch == 1 || ch == 3)
        c = 3u - c;
    return c + 1u;
}

// A chain whose interval holds at most cb.vmax rows at a depth where nothing has
// been recorded yet (o <= t) is handed to k_verify, which finishes it on the
// text; returns true when it was (the backward search of the chain stops).
template <typename R>
__device__ __forceinline__ bool hand_over(const ChainBuf &cb, uint64_t chain, uint32_t o,
                                          uint32_t t, R l, uint32_t size)
{
    if (size > cb.vmax || o > t)
        return false;
    cb.vl[chain] = (uint64_t) l;
    cb.vjob[chain] = size | (o << 16);
    return true;
}

// One chain per lane, four adjacent lanes per read.
template <typename R>
__global__ void __launch_bounds__(256) k_find1(Geom G, DirViewT<R> fw, DirViewT<R> bw, uint64_t r0,
                                              uint64_t nr, uint32_t t, CTab<R> C,
                                              const SeedT<R> *seeds, uint32_t kseed,
                                              ChainBuf out, unsigned long long *stats)
{
    uint64_t tid = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t rl = tid >> 2;
    int ch = (int) (tid & 3);
    uint32_t steps = 0, lines = 0;
    if (rl < nr) {
    uint64_t r = r0 + rl;
    uint64_t s = pq_read_start(G, r);
    uint32_t L = pq_read_len(G, r);
    const bool fwsearch = (ch == 0 || ch == 3);
    const OccLine *occ = fwsearch ? fw.occ : bw.occ;
    const uint32_t minsize = (ch == 0 || ch == 2) ? 2u : 1u;
    uint32_t o, size;
    R l;
    SeedT<R> sd = {0, SEED_NONE};
    if (seeds && kseed > 0 && kseed < L)
        sd = seeds[rl * 4 + ch];
    // a chain the seed kernel finished keeps its records
    const bool done = sd.size == SEED_DONE;
    if (sd.size != SEED_NONE) {
        o = kseed;
        l = sd.l;
        size = done ? 0u : sd.size;
    } else {
        // C[0..5] = 0, 0 (A), C[2], C[3], C[4], C[5] (index.c:367-372)
        unsigned c = chain_sym(fw.text2, s, L, ch, 1);
        o = 1;
        l = pq_pick6(C.c, c);
        size = (uint32_t) (pq_pick6(C.c, c + 1) - l);
    }
    uint32_t nout = 0;
    uint64_t base = (rl * 4 + (uint64_t) ch) * out.cap;
    out.vjob[rl * 4 + ch] = 0;
    while (o < L && size >= minsize) {
        if (hand_over(out, rl * 4 + ch, o, t, l, size))
            break;
        // the next pattern symbol is read with the interval's lines
        const bool more = o + 1 < L;
        const unsigned c = more ? chain_sym(fw.text2, s, L, ch, o + 1) : 1u;
        const uint64_t u1 = (uint64_t) l + size;
        const uint64_t wl = (uint64_t) l >> 7, wu = u1 >> 7;
        OccLine ln;
        pq_load_line(occ, wl, ln);
        const StepRanks a = pq_step_ranks(ln, wl, (unsigned) (l & 127), c);
        if (wu != wl)
            pq_load_line(occ, wu, ln);
        const StepRanks b = pq_step_ranks(ln, wu, (unsigned) (u1 & 127), c);
        steps++;
        lines += wl == wu ? 1u : 2u;
        if (o >= t) {
            uint32_t cnt = b.d - a.d;
            if (cnt && nout < out.cap) {
                out.lc[base + nout] = make_uint2(a.d, cnt);
                out.o[base + nout] = (uint16_t) o;
                nout++;
            }
        }
        if (!more)
            break;
        l = (fwsearch ? pq_pick5(fw.cf, c) : pq_pick5(bw.cf, c)) + a.c;
        size = b.c - a.c;
        o++;
    }
    if (!done)
        out.n[rl * 4 + ch] = nout;
    }
    if (stats) {
        uint64_t s2 = steps, ln = lines;
        for (int d = 32; d > 0; d >>= 1) {
            s2 += __shfl_xor(s2, d, 64);
            ln += __shfl_xor(ln, d, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            wstat_add(stats, WS_POSITIONS, 2 * s2);
            wstat_add(stats, WS_LINES, ln);
        }
    }
}

// Lane I of the quad (four adjacent lanes) broadcast to the whole quad.
template <int I> __device__ __forceinline__ uint32_t quad_bcast(uint32_t v)
{
    return (uint32_t) __builtin_amdgcn_mov_dpp((int) v, I * 0x55, 0xF, 0xF, false);
}

template <int I> __device__ __forceinline__ uint64_t quad_bcast64(uint64_t v)
{
    return ((uint64_t) quad_bcast<I>((uint32_t) (v >> 32)) << 32) | quad_bcast<I>((uint32_t) v);
}

// Lane q of a quad reads bytes [16q, 16q+16) of the line of quad lane I, when
// that lane needs one (quad-uniform): one 64-byte access per line.
template <int I>
__device__ __forceinline__ void quad_fetch(const OccLine *occ, uint64_t w, bool need, uint4 *xq,
                                           unsigned q)
{
    const uint64_t wi = quad_bcast64<I>(w);
    if (quad_bcast<I>(need ? 1u : 0u))
        xq[I * 4 + q] = reinterpret_cast<const uint4 *>(occ + wi)[q];
}

// Line of this lane from the quarters its quad stored in xq (lane-major).
__device__ __forceinline__ void quad_line(const uint4 *xq, unsigned q, OccLine &ln)
{
    const uint4 q0 = xq[q * 4 + 0], q1 = xq[q * 4 + 1], q2 = xq[q * 4 + 2], q3 = xq[q * 4 + 3];
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

// One chain per lane as k_find1, with the occurrence lines of a read's four
// chains loaded by its four lanes together: lane q fetches quarter q of each
// lane's line, one 64-byte access per line instead of four 16-byte accesses
// from one lane, and the quarters are exchanged through LDS.  Where the lines
// span more memory than the address translation caches cover (gigabytes of
// lines), every access pays a translation: a quarter of the accesses then
// doubles the rate of scattered line loads.  Chains 0 and 3 search the forward
// index, 1 and 2 the backward one, so the index of each quad lane is known at
// compile time.  The four lanes of a read iterate while any of them is live.
template <typename R>
__global__ void __launch_bounds__(256) k_find1q(Geom G, DirViewT<R> fw, DirViewT<R> bw,
                                               uint64_t r0, uint64_t nr, uint32_t t, CTab<R> C,
                                               const SeedT<R> *seeds, uint32_t kseed,
                                               ChainBuf out, unsigned long long *stats)
{
    __shared__ uint4 xch[256 * 4];
    uint64_t tid = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t rl = tid >> 2;
    const int ch = (int) (tid & 3);
    const unsigned q = threadIdx.x & 3;
    uint4 *xq = xch + (threadIdx.x & ~3u) * 4;
    uint32_t steps = 0, lines = 0;
    const bool valid = rl < nr;
    const uint64_t r = r0 + (valid ? rl : 0);
    const uint64_t s = pq_read_start(G, r);
    const uint32_t L = pq_read_len(G, r);
    const bool fwsearch = (ch == 0 || ch == 3);
    const uint32_t minsize = (ch == 0 || ch == 2) ? 2u : 1u;
    uint32_t o = 1, size = 0;
    R l = 0;
    SeedT<R> sd = {0, SEED_NONE};
    if (valid && seeds && kseed > 0 && kseed < L)
        sd = seeds[rl * 4 + ch];
    const bool done = sd.size == SEED_DONE;   // finished by the seed kernel
    if (sd.size != SEED_NONE) {
        o = kseed;
        l = sd.l;
        size = done ? 0u : sd.size;
    } else if (valid) {
        unsigned c = chain_sym(fw.text2, s, L, ch, 1);
        l = pq_pick6(C.c, c);
        size = (uint32_t) (pq_pick6(C.c, c + 1) - l);
    }
    uint32_t nout = 0;
    const uint64_t base = (rl * 4 + (uint64_t) ch) * out.cap;
    if (valid)
        out.vjob[rl * 4 + ch] = 0;
    bool live = valid && o < L && size >= minsize;
    if (live && hand_over(out, rl * 4 + ch, o, t, l, size))
        live = false;
    const unsigned qshift = threadIdx.x & 60u;
    while ((__ballot(live) >> qshift) & 0xFull) {
        const bool more = live && o + 1 < L;
        const unsigned c = more ? chain_sym(fw.text2, s, L, ch, o + 1) : 1u;
        const uint64_t u1 = (uint64_t) l + size;
        const uint64_t wl = (uint64_t) l >> 7, wu = u1 >> 7;
        const bool two = live && wu != wl;
        quad_fetch<0>(fw.occ, wl, live, xq, q);
        quad_fetch<1>(bw.occ, wl, live, xq, q);
        quad_fetch<2>(bw.occ, wl, live, xq, q);
        quad_fetch<3>(fw.occ, wl, live, xq, q);
        wave_sync();
        OccLine ln;
        quad_line(xq, q, ln);
        const StepRanks a = pq_step_ranks(ln, wl, (unsigned) (l & 127), c);
        wave_sync();
        if ((__ballot(two) >> qshift) & 0xFull) {
            quad_fetch<0>(fw.occ, wu, two, xq, q);
            quad_fetch<1>(bw.occ, wu, two, xq, q);
            quad_fetch<2>(bw.occ, wu, two, xq, q);
            quad_fetch<3>(fw.occ, wu, two, xq, q);
            wave_sync();
            if (two)
                quad_line(xq, q, ln);
            wave_sync();
        }
        const StepRanks b = pq_step_ranks(ln, wu, (unsigned) (u1 & 127), c);
        if (live) {
            steps++;
            lines += two ? 2u : 1u;
            if (o >= t) {
                uint32_t cnt = b.d - a.d;
                if (cnt && nout < out.cap) {
                    out.lc[base + nout] = make_uint2(a.d, cnt);
                    out.o[base + nout] = (uint16_t) o;
                    nout++;
                }
            }
            if (more) {
                l = (fwsearch ? pq_pick5(fw.cf, c) : pq_pick5(bw.cf, c)) + a.c;
                size = b.c - a.c;
                o++;
                live = size >= minsize && !hand_over(out, rl * 4 + ch, o, t, l, size);
            } else {
                live = false;
            }
        }
    }
    if (valid && !done)
        out.n[rl * 4 + ch] = nout;
    if (stats) {
        uint64_t s2 = steps, ln = lines;
        for (int d = 32; d > 0; d >>= 1) {
            s2 += __shfl_xor(s2, d, 64);
            ln += __shfl_xor(ln, d, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            wstat_add(stats, WS_POSITIONS, 2 * s2);
            wstat_add(stats, WS_LINES, ln);
        }
    }
}

// Two chains per lane: FB and BB search the forward index, FF and BF the
// backward one, so a lane takes the two chains of one index of a read (same
// occurrence lines and step bases) and keeps one backward step of each in
// flight: two independent line loads where one chain has only one.
template <typename R>
__global__ void __launch_bounds__(256) k_find2(Geom G, DirViewT<R> fw, DirViewT<R> bw,
                                               uint64_t r0, uint64_t nr, uint32_t t, CTab<R> C,
                                               const SeedT<R> *seeds, uint32_t kseed,
                                               ChainBuf out, unsigned long long *stats)
{
    const uint64_t tid = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t rl = tid >> 1;
    const bool bwi = (tid & 1) != 0;
    uint32_t steps = 0, lines = 0;
    if (rl < nr) {
        const uint64_t r = r0 + rl;
        const uint64_t s = pq_read_start(G, r);
        const uint32_t L = pq_read_len(G, r);
        const OccLine *occ = bwi ? bw.occ : fw.occ;
        // the lane's step bases as an array (the compiler keeps it in LDS): with
        // selects from the kernel arguments instead, the seeded search of the
        // 100 bp workload took 6% longer per launch
        R cf[5];
#pragma unroll
        for (int q = 0; q < 5; q++)
            cf[q] = bwi ? bw.cf[q] : fw.cf[q];
        // chains (FB, BB) on the forward index, (FF, BF) on the backward one
        const int ch[2] = {bwi ? 1 : 0, bwi ? 2 : 3};
        const uint32_t minsize[2] = {bwi ? 1u : 2u, bwi ? 2u : 1u};
        uint32_t o[2], size[2], nout[2];
        R l[2];
        bool live[2], done[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            SeedT<R> sd = {0, SEED_NONE};
            if (seeds && kseed > 0 && kseed < L)
                sd = seeds[rl * 4 + ch[q]];
            done[q] = sd.size == SEED_DONE;   // finished by the seed kernel
            if (sd.size != SEED_NONE) {
                o[q] = kseed;
                l[q] = sd.l;
                size[q] = done[q] ? 0u : sd.size;
            } else {
                // C[0..5] = 0, 0 (A), C[2], C[3], C[4], C[5] (index.c:367-372)
                unsigned c = chain_sym(fw.text2, s, L, ch[q], 1);
                o[q] = 1;
                l[q] = pq_pick6(C.c, c);
                size[q] = (uint32_t) (pq_pick6(C.c, c + 1) - l[q]);
            }
            nout[q] = 0;
            out.vjob[rl * 4 + ch[q]] = 0;
            live[q] = o[q] < L && size[q] >= minsize[q] &&
                      !hand_over(out, rl * 4 + ch[q], o[q], t, l[q], size[q]);
        }
        while (live[0] || live[1]) {
            bool more[2];
            unsigned c[2];
            uint64_t wl[2], wu[2], u1[2];
            OccLine la[2], lb[2];
            // both chains' lines (and next symbols) requested before either is used
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (!live[q])
                    continue;
                more[q] = o[q] + 1 < L;
                c[q] = more[q] ? chain_sym(fw.text2, s, L, ch[q], o[q] + 1) : 1u;
                u1[q] = (uint64_t) l[q] + size[q];
                wl[q] = (uint64_t) l[q] >> 7;
                wu[q] = u1[q] >> 7;
                pq_load_line(occ, wl[q], la[q]);
                if (wu[q] != wl[q])
                    pq_load_line(occ, wu[q], lb[q]);
                else
                    lb[q] = la[q];
            }
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (!live[q])
                    continue;
                const StepRanks a = pq_step_ranks(la[q], wl[q], (unsigned) (l[q] & 127), c[q]);
                const StepRanks b = pq_step_ranks(lb[q], wu[q], (unsigned) (u1[q] & 127), c[q]);
                steps++;
                lines += wl[q] == wu[q] ? 1u : 2u;
                if (o[q] >= t) {
                    const uint32_t cnt = b.d - a.d;
                    if (cnt && nout[q] < out.cap) {
                        const uint64_t at = (rl * 4 + (uint64_t) ch[q]) * out.cap + nout[q];
                        out.lc[at] = make_uint2(a.d, cnt);
                        out.o[at] = (uint16_t) o[q];
                        nout[q]++;
                    }
                }
                if (!more[q]) {
                    live[q] = false;
                    continue;
                }
                l[q] = pq_pick5(cf, c[q]) + a.c;
                size[q] = b.c - a.c;
                o[q]++;
                live[q] = o[q] < L && size[q] >= minsize[q] &&
                          !hand_over(out, rl * 4 + ch[q], o[q], t, l[q], size[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (!done[q])
                out.n[rl * 4 + ch[q]] = nout[q];
    }
    if (stats) {
        uint64_t s2 = steps, ln = lines;
        for (int d = 32; d > 0; d >>= 1) {
            s2 += __shfl_xor(s2, d, 64);
            ln += __shfl_xor(ln, d, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            wstat_add(stats, WS_POSITIONS, 2 * s2);
            wstat_add(stats, WS_LINES, ln);
        }
    }
}

// The seeded search's chains, from the work list the seed kernel wrote (only the
// chains it left unfinished: most are finished or empty at the seed depth).  One
// chain per lane; a lane whose chain ends takes the next listed chain, so the
// lanes of a wave stay busy until the list is drained instead of idling behind
// the wave's longest chain.  Idle lanes take chains together, one atomic per
// wave, once a quarter of the wave is idle or none of it is busy.  Launched as a
// grid that fills the GPU once.
template <typename R>
__global__ void __launch_bounds__(256) k_find_list(Geom G, DirViewT<R> fw, DirViewT<R> bw,
                                                   uint64_t r0, uint32_t nchains, uint32_t t,
                                                   CTab<R> C, const SeedT<R> *seeds,
                                                   uint32_t kseed, ChainBuf out,
                                                   unsigned long long *stats)
{
    // without a list (the plain search) every chain of the batch is listed in order
    const uint32_t total = out.wq ? out.wq_cnt[0] : nchains;
    const int lane = threadIdx.x & 63;
    uint32_t steps = 0, lines = 0;
    bool live = false, more = true;
    uint64_t chain = 0, s = 0;
    uint32_t L = 0, o = 0, size = 0, nout = 0, minsize = 1;
    int ch = 0;
    R l = 0;
    for (;;) {
        const bool want = !live && more;
        const uint64_t wm = __ballot(want);
        if (wm && (__popcll(wm) >= 16 || !__ballot(live))) {
            const int lead = __ffsll((long long) wm) - 1;
            uint32_t base = 0;
            if (lane == lead)
                base = atomicAdd(&out.wq_cnt[1], (uint32_t) __popcll(wm));
            base = __shfl(base, lead, 64);
            if (want) {
                const uint32_t idx = base + (uint32_t) __popcll(wm & ((1ull << lane) - 1));
                if (idx < total) {
                    chain = out.wq ? out.wq[idx] : idx;
                    const uint64_t r = r0 + (chain >> 2);
                    ch = (int) (chain & 3);
                    s = pq_read_start(G, r);
                    L = pq_read_len(G, r);
                    minsize = (ch == 0 || ch == 2) ? 2u : 1u;
                    SeedT<R> sd = {0, SEED_NONE};
                    if (seeds && kseed > 0 && kseed < L)
                        sd = seeds[chain];
                    if (sd.size != SEED_NONE) {
                        o = kseed;
                        l = sd.l;
                        size = sd.size;
                    } else {
                        // C[0..5] = 0, 0 (A), C[2], C[3], C[4], C[5] (index.c:367-372)
                        const unsigned c = chain_sym(fw.text2, s, L, ch, 1);
                        o = 1;
                        l = pq_pick6(C.c, c);
                        size = (uint32_t) (pq_pick6(C.c, c + 1) - l);
                    }
                    nout = 0;
                    if (!out.wq)
                        out.vjob[chain] = 0;
                    live = o < L && size >= minsize && !hand_over(out, chain, o, t, l, size);
                    if (!live)
                        out.n[chain] = 0;
                } else {
                    more = false;
                }
            }
        }
        if (!__ballot(live)) {
            if (!__ballot(more))
                break;
            continue;
        }
        if (live) {
            const bool fwsearch = ch == 0 || ch == 3;
            const bool next = o + 1 < L;
            const unsigned c = next ? chain_sym(fw.text2, s, L, ch, o + 1) : 1u;
            const uint64_t u1 = (uint64_t) l + size;
            const uint64_t wl = (uint64_t) l >> 7, wu = u1 >> 7;
            const OccLine *occ = fwsearch ? fw.occ : bw.occ;
            OccLine la, lb;
            pq_load_line(occ, wl, la);
            if (wu != wl)
                pq_load_line(occ, wu, lb);
            else
                lb = la;
            const StepRanks a = pq_step_ranks(la, wl, (unsigned) (l & 127), c);
            const StepRanks b = pq_step_ranks(lb, wu, (unsigned) (u1 & 127), c);
            steps++;
            lines += wl == wu ? 1u : 2u;
            if (o >= t) {
                const uint32_t cnt = b.d - a.d;
                if (cnt && nout < out.cap) {
                    out.lc[chain * out.cap + nout] = make_uint2(a.d, cnt);
                    out.o[chain * out.cap + nout] = (uint16_t) o;
                    nout++;
                }
            }
            if (next) {
                l = (fwsearch ? pq_pick5(fw.cf, c) : pq_pick5(bw.cf, c)) + a.c;
                size = b.c - a.c;
                o++;
                live = size >= minsize && !hand_over(out, chain, o, t, l, size);
            } else {
                live = false;
            }
            if (!live)
                out.n[chain] = nout;
        }
    }
    if (stats) {
        uint64_t s2 = steps, ln = lines;
        for (int d = 32; d > 0; d >>= 1) {
            s2 += __shfl_xor(s2, d, 64);
            ln += __shfl_xor(ln, d, 64);
        }
        if (lane == 0) {
            wstat_add(stats, WS_POSITIONS, 2 * s2);
            wstat_add(stats, WS_LINES, ln);
        }
    }
}

// ---- verification ----------------------------------------------------------------------

// Chains handed over by the search kernels (cb.vjob): one lane per chain
// finishes it on the text (verify_chain).
template <typename R>
__global__ void __launch_bounds__(256) k_verify(Geom G, DirViewT<R> fw, DirViewT<R> bw,
                                               ChainBuf cb, uint64_t r0, uint64_t nr, uint32_t t,
                                               unsigned long long *stats)
{
    const uint64_t chain = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t rl = chain >> 2;
    const int ch = (int) (chain & 3);
    const uint32_t job = rl < nr ? cb.vjob[chain] : 0u;
    uint32_t verified = 0;
    __shared__ uint64_t stg[STAGE_SLOTS * STAGE_STRIDE];   // candidate staging
    if (job) {
        const uint32_t size = job & 0xFFFFu, o0 = job >> 16;
        const uint64_t r = r0 + rl;
        const bool fwsearch = ch == 0 || ch == 3;
        cb.n[chain] = verify_chain<R>(
            G, fwsearch ? fw.sa : bw.sa, fwsearch ? bw.text2 : fw.text2,
            ch < 2 ? bw.text2 : fw.text2, (ch == 1 || ch == 3) ? ~0ull : 0ull,
            pq_read_start(G, r), pq_read_len(G, r), o0, cb.vl[chain], size,
            (ch == 0 || ch == 2) ? 2u : 1u, t, cb.lc + chain * cb.cap, cb.o + chain * cb.cap,
            stg + threadIdx.x);
        verified = size;
    }
    if (stats) {
        uint64_t v = verified;
        for (int d = 32; d > 0; d >>= 1)
            v += __shfl_xor(v, d, 64);
        if ((threadIdx.x & 63) == 0 && v)
            wstat_add(stats, WS_VERIFY_ROWS, v);
    }
}

// ---- extraction ----------------------------------------------------------------------------

struct ItemMeta {
    uint32_t m;
    uint16_t o;
    uint16_t klen;
    uint8_t rank, type, valid, pad;
};

// -1/+1: a sorts before/after b with a difference inside the common length;
// -2: a is a proper prefix of b; +2: b is a proper prefix of a; 0: equal.
__device__ __forceinline__ int key_cmp(const uint64_t *a, uint32_t la, const uint64_t *b,
                                       uint32_t lb)
{
    uint32_t mn = la < lb ? la : lb;
    uint32_t full = mn >> 5, rem = mn & 31;
    for (uint32_t w = 0; w < full; w++) {
        uint64_t x = a[w], y = b[w];
        if (x != y)
            return x < y ? -1 : 1;
    }
    if (rem) {
        uint64_t mask = ~0ull << (64 - 2 * rem);
        uint64_t x = a[full] & mask, y = b[full] & mask;
        if (x != y)
            return x < y ? -1 : 1;
    }
    return la < lb ? -2 : (la > lb ? 2 : 0);
}

// Leading codes two keys share (at most the shorter length).
__device__ __forceinline__ uint32_t key_lcp(const uint64_t *a, uint32_t la, const uint64_t *b,
                                            uint32_t lb)
{
    const uint32_t mn = la < lb ? la : lb;
    for (uint32_t w = 0; 32 * w < mn; w++) {
        uint64_t x = a[w] ^ b[w];
        if (x) {
            uint32_t l = 32 * w + (uint32_t) __clzll((long long) x) / 2;
            return l < mn ? l : mn;
        }
    }
    return mn;
}

#define EDGE_SLOT 16          // edges per read side staged without any atomic

struct ExtOut {
    int32_t *ovB, *ovF;
    uint32_t *numB, *numF;
    uint64_t *recB, *recF;   // pool position of a side's edges when it has more
                             // than EDGE_SLOT of them
    uint64_t *staged;        // EDGE_SLOT edges per (read, side)
    uint64_t *edges;         // pool for larger sides
    uint64_t edge_cap;
    unsigned long long *edge_count;
    uint32_t *overflow_reads;     // reads left for the global-memory pass
    uint32_t *n_overflow;
    uint32_t *flags;              // bit0: edge buffer full, bit1: item limit
    unsigned long long *xstats;   // work counters (WS_ITEMS, WS_TIER_READS + tier)
    uint32_t tier;                // 0..2: LDS tiers in launch order, 3: global scratch
};

// Items and reads the groups of a wave finished, added to the work counters.
template <int GW>
__device__ __forceinline__ void ext_count(const ExtOut &out, uint64_t items, uint64_t reads)
{
    if ((threadIdx.x & (GW - 1)) != 0)
        items = reads = 0;
    for (int d = 32; d > 0; d >>= 1) {
        items += __shfl_xor(items, d, 64);
        reads += __shfl_xor(reads, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        wstat_add(out.xstats, WS_ITEMS, items);
        wstat_add(out.xstats, WS_TIER_READS + out.tier, reads);
    }
}


// Lanes of this group of GW lanes set in a wave-wide ballot, at the group's
// lane positions.
template <int GW> __device__ __forceinline__ uint64_t group_bits(bool x)
{
    uint64_t b = __ballot(x);
    if (GW == 64)
        return b;
    const int shift = (int) (threadIdx.x % 64) & ~(GW - 1);
    return (b >> shift) & ((1ull << GW) - 1ull);
}

// Strict order of the extraction: valid items by (key, rank, o, m), then
// invalid ones, then the padding indices >= k.
__device__ __forceinline__ bool item_less(uint32_t a, uint32_t b, uint32_t k, uint32_t kw,
                                          const uint64_t *keys, const ItemMeta *meta)
{
    if (a >= k || b >= k)
        return a < b;
    const ItemMeta ma = meta[a], mb = meta[b];
    if (ma.valid != mb.valid)
        return ma.valid > mb.valid;
    if (!ma.valid)
        return a < b;
    int c = key_cmp(keys + (uint64_t) a * kw, ma.klen, keys + (uint64_t) b * kw, mb.klen);
    if (c != 0)
        return c < 0;
    if (ma.rank != mb.rank)
        return ma.rank < mb.rank;
    if (ma.o != mb.o)
        return ma.o < mb.o;
    return ma.m < mb.m;
}

// Edge emission of one read side evaluated by one group of GW lanes on items
// sorted by (key, rank, o, m), the order the edges are emitted in.  The valid
// items sort first; with lcp[i] the codes shared by sorted items i-1 and i:
//  - a key has another key as a proper prefix iff the chain of keys before it
//    that are prefixes of their successors reaches a shorter one: with
//    M = the shortest key of the chain ending at the previous item, the chain
//    continues iff M <= lcp[i], and the item is dropped iff also M < len[i];
//  - among equal keys the first group by (rank, o) is kept (a block of equal
//    keys starts where lcp[i] is below either length);
//  - the read itself is never emitted;
// the rules of the reference's depth-first walk (string_graph.c:386-731) as
// restated in the header.  perm[0..P) holds the sorted item indices (P a power
// of two >= k); emitf and lcpv are indexed by sorted position.
template <int GW>
__device__ void extract_emit(uint64_t r, int side, uint32_t kw, uint32_t k, uint32_t P,
                             const uint64_t *keys, const ItemMeta *meta, uint8_t *emitf,
                             const uint32_t *perm, uint16_t *lcpv, const ExtOut &out);

// Sort and emission of one read side by one group of GW lanes: a bitonic
// network over the item indices (P: a power of two >= k, the same for every
// group of a wave -- the tier's item limit when several groups share a wave --
// so that all loops have the same trip count in every group).  perm reuses the
// entry prefix array.
template <int GW>
__device__ void extract_sorted(uint64_t r, int side, uint32_t kw, uint32_t k, uint32_t P,
                               const uint64_t *keys, const ItemMeta *meta, uint8_t *emitf,
                               uint32_t *perm, uint16_t *lcpv, const ExtOut &out)
{
    const int lane = threadIdx.x & (GW - 1);
    for (uint32_t i = lane; i < P; i += GW)
        perm[i] = i;
    wave_sync();
    for (uint32_t s2 = 2; s2 <= P; s2 <<= 1) {
        for (uint32_t j2 = s2 >> 1; j2 > 0; j2 >>= 1) {
            for (uint32_t p = lane; p < P / 2; p += GW) {
                const uint32_t i = 2 * j2 * (p / j2) + (p % j2), ip = i + j2;
                const bool up = (i & s2) == 0;
                const uint32_t x = perm[i], y = perm[ip];
                if (item_less(y, x, k, kw, keys, meta) == up) {
                    perm[i] = y;
                    perm[ip] = x;
                }
            }
            wave_sync();
        }
    }
    extract_emit<GW>(r, side, kw, k, P, keys, meta, emitf, perm, lcpv, out);
}

template <int GW>
__device__ void extract_emit(uint64_t r, int side, uint32_t kw, uint32_t k, uint32_t P,
                             const uint64_t *keys, const ItemMeta *meta, uint8_t *emitf,
                             const uint32_t *perm, uint16_t *lcpv, const ExtOut &out)
{
    const int lane = threadIdx.x & (GW - 1);
    uint32_t nv = 0;
    for (uint32_t q0 = 0; q0 < P; q0 += GW) {
        const uint32_t q = q0 + lane;
        nv += (uint32_t) __popcll(group_bits<GW>(q < k && meta[q].valid));
    }
    for (uint32_t i = lane; i < P; i += GW) {
        if (i < nv) {
            uint32_t l = 0;
            if (i > 0) {
                const uint32_t a = perm[i - 1], b = perm[i];
                l = key_lcp(keys + (uint64_t) a * kw, meta[a].klen, keys + (uint64_t) b * kw,
                            meta[b].klen);
            }
            lcpv[i] = (uint16_t) l;
        }
    }
    wave_sync();
    // the ordered pass over the sorted items with the state in registers
    // (group-uniform)
    uint32_t M = 0, prevlen = 0, bf_rank = 0, bf_o = 0, emitted = 0;
    if constexpr (GW == 64) {
        // 64 sorted items per round, one per lane, read back lane by lane
        for (uint32_t c0 = 0; c0 < nv; c0 += 64) {
            const uint32_t i = c0 + lane;
            uint32_t li = 0, ai = 0, rk = 0, oi = 0, self = 0;
            if (i < nv) {
                const ItemMeta mt = meta[perm[i]];
                li = mt.klen;
                ai = lcpv[i];
                rk = mt.rank;
                oi = mt.o;
                self = mt.m == (uint32_t) r ? 1u : 0u;
            }
            uint32_t keepv = 0;
            const uint32_t nt = nv - c0 < 64 ? nv - c0 : 64;
            for (uint32_t t = 0; t < nt; t++) {
                const uint32_t Lt = __builtin_amdgcn_readlane(li, t);
                const uint32_t At = __builtin_amdgcn_readlane(ai, t);
                const uint32_t Rt = __builtin_amdgcn_readlane(rk, t);
                const uint32_t Ot = __builtin_amdgcn_readlane(oi, t);
                const uint32_t St = __builtin_amdgcn_readlane(self, t);
                const bool first = c0 + t == 0;
                const bool chain = !first && M <= At;
                const bool dropped = chain && M < Lt;
                const bool same_key = !first && At == Lt && At == prevlen;
                if (!same_key) {
                    bf_rank = Rt;
                    bf_o = Ot;
                }
                const bool keep = !dropped && Rt == bf_rank && Ot == bf_o && !St;
                M = chain ? (M < Lt ? M : Lt) : Lt;
                prevlen = Lt;
                emitted += keep ? 1u : 0u;
                keepv = (uint32_t) lane == t ? (keep ? 1u : 0u) : keepv;
            }
            if (i < nv)
                emitf[i] = (uint8_t) keepv;
        }
    } else {
        // a few items per group: walk them in order, every lane of the group
        // reading the same LDS entries
        for (uint32_t i = 0; i < nv; i++) {
            {
                const ItemMeta mt = meta[perm[i]];
                const uint32_t Lt = mt.klen, At = lcpv[i];
                const bool first = i == 0;
                const bool chain = !first && M <= At;
                const bool dropped = chain && M < Lt;
                const bool same_key = !first && At == Lt && At == prevlen;
                if (!same_key) {
                    bf_rank = mt.rank;
                    bf_o = mt.o;
                }
                const bool keep =
                    !dropped && mt.rank == bf_rank && mt.o == bf_o && mt.m != (uint32_t) r;
                M = chain ? (M < Lt ? M : Lt) : Lt;
                prevlen = Lt;
                emitted += keep ? 1u : 0u;
                if (lane == 0)
                    emitf[i] = keep ? 1u : 0u;
            }
        }
    }
    wave_sync();
    uint64_t base = 0;
    uint64_t *dst = out.staged + (r * 2 + side) * EDGE_SLOT;
    if (lane == 0) {
        if (emitted > EDGE_SLOT) {
            base = atomicAdd(out.edge_count, (unsigned long long) emitted);
            if (base + emitted > out.edge_cap)
                atomicOr(out.flags, 1u);
        }
        if (side) {
            out.ovF[r] = (int32_t) nv;
            out.numF[r] = emitted;
            out.recF[r] = base;
        } else {
            out.ovB[r] = (int32_t) nv;
            out.numB[r] = emitted;
            out.recB[r] = base;
        }
    }
    base = __shfl(base, 0, GW);
    uint64_t cap = EDGE_SLOT;
    if (emitted > EDGE_SLOT) {
        dst = out.edges + base;
        cap = base < out.edge_cap ? out.edge_cap - base : 0;
    }
    // output position = number of kept items before it in sorted order
    uint32_t carry = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t c0 = 0; c0 < P; c0 += GW) {
        const uint32_t i = c0 + lane;
        const bool kf = i < nv && emitf[i];
        const uint64_t bal = group_bits<GW>(kf);
        const uint32_t pos = carry + (uint32_t) __popcll(bal & lt);
        if (kf && pos < cap) {
            const ItemMeta mi = meta[perm[i]];
            dst[pos] = ((uint64_t) mi.m << 32) | ((uint64_t) mi.o << 8) | mi.type;
        }
        carry += (uint32_t) __popcll(bal);
    }
    wave_sync();
}

// Item (type of chain ch, overlap o, read m) with its extension key: FB/BB
// m[o..) from the forward text; FF/BF rc(m)[o..) as the complement of the
// backward text; a read contained in the overlap is no candidate (invalid).
template <typename R>
__device__ __forceinline__ void item_fill(const Geom &G, const DirViewT<R> &fw,
                                          const DirViewT<R> &bw, int ch, uint32_t rank,
                                          uint32_t o, uint64_t m, uint32_t kw, uint64_t *kq,
                                          ItemMeta *meta)
{
    const uint32_t Lm = pq_read_len(G, m);
    ItemMeta mt;
    mt.m = (uint32_t) m;
    mt.o = (uint16_t) o;
    mt.valid = o < Lm;
    mt.klen = (uint16_t) (mt.valid ? Lm - o : 0);
    mt.rank = (uint8_t) rank;
    mt.type = ch == 0 ? 1 : ch == 1 ? 2 : ch == 2 ? 0 : 3;   // FB FF BF BB
    mt.pad = 0;
    const bool plain = (ch == 0 || ch == 3);
    const uint64_t *tx = plain ? fw.text2 : bw.text2;
    const uint64_t ks = pq_read_start(G, m) + o;
    for (uint32_t w = 0; w < kw; w++) {
        uint64_t x = 0;
        if (32 * w < mt.klen) {
            x = pq_window(tx, ks + 32 * w);
            if (!plain)
                x = ~x;
            const uint32_t left = mt.klen - 32 * w;
            if (left < 32)
                x &= ~0ull << (64 - 2 * left);
        }
        kq[w] = x;
    }
    *meta = mt;
}

// Process one read with one group of GW lanes (a wave or a quarter wave; the
// groups of a wave work on different reads).  keys/meta/emitf/pre: this group's
// storage (LDS or global) for kmax items.  Returns false, writing nothing, if a
// side has more than kmax items.
template <int GW, typename R>
__device__ bool extract_read(const Geom &G, const DirViewT<R> &fw, const DirViewT<R> &bw,
                             const ChainBuf &cb, uint64_t rl, uint64_t r, uint32_t kw,
                             uint64_t *keys, ItemMeta *meta, uint8_t *emitf, uint32_t *pre,
                             uint16_t *lcpv, uint32_t kmax, const ExtOut &out, uint32_t &items)
{
    const int lane = threadIdx.x & (GW - 1);
    uint32_t cnt4[4];
    items = 0;
    for (int c = 0; c < 4; c++)
        cnt4[c] = cb.n[rl * 4 + c];
    // items per side first, so that an oversized read is left untouched
    for (int side = 0; side < 2; side++) {
        const int chA = side ? 0 : 2, chB = side ? 1 : 3;
        uint32_t k = 0;
        for (uint32_t e = lane; e < cnt4[chA] + cnt4[chB]; e += GW) {
            uint64_t at = (e < cnt4[chA]) ? (rl * 4 + chA) * cb.cap + e
                                          : (rl * 4 + chB) * cb.cap + (e - cnt4[chA]);
            k += entry_count(cb.lc[at]);
        }
        for (int d = GW / 2; d > 0; d >>= 1)
            k += __shfl_xor(k, d, GW);
        if (k > kmax)
            return false;
    }
    // side 0 = B (chains 2: BF, 3: BB), side 1 = F (chains 0: FB, 1: FF)
    for (int side = 0; side < 2; side++) {
        const int chA = side ? 0 : 2, chB = side ? 1 : 3;
        const uint32_t nA = cnt4[chA], nB = cnt4[chB], ne = nA + nB;
        // exclusive prefix of the counts over the side's entries
        uint32_t carry = 0;
        for (uint32_t e0 = 0; e0 < ne; e0 += GW) {
            uint32_t e = e0 + lane;
            uint32_t v = 0;
            if (e < ne) {
                uint64_t at = (e < nA) ? (rl * 4 + chA) * cb.cap + e
                                       : (rl * 4 + chB) * cb.cap + (e - nA);
                v = entry_count(cb.lc[at]);
            }
            uint32_t inc = v;
            for (int d = 1; d < GW; d <<= 1) {
                uint32_t x = __shfl_up(inc, d, GW);
                if (lane >= d)
                    inc += x;
            }
            if (e < ne)
                pre[e] = carry + inc - v;
            carry += __shfl(inc, GW - 1, GW);
        }
        const uint32_t k = carry;
        items += k;
        wave_sync();
        // expand the candidate intervals into items with their extension keys:
        // two items per lane at a time, so that each lane has two independent
        // chains of loads (candidate record -> reads_index -> key words) in flight
        auto entry_of = [&](uint32_t q) -> uint32_t {   // e with pre[e] <= q < pre[e+1]
            uint32_t lo = 0, hi = ne;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (pre[mid] <= q)
                    lo = mid;
                else
                    hi = mid;
            }
            return lo;
        };
        for (uint32_t q0 = lane; q0 < k; q0 += 2 * GW) {
            uint32_t qq[2] = {q0, q0 + GW};
            bool on[2] = {true, q0 + GW < k};
            uint32_t ee[2], oo[2];
            bool firsts[2];
            uint2 lcs[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                ee[h] = on[h] ? entry_of(qq[h]) : 0u;
                firsts[h] = ee[h] < nA;
                const uint64_t at = firsts[h] ? (rl * 4 + chA) * cb.cap + ee[h]
                                              : (rl * 4 + chB) * cb.cap + (ee[h] - nA);
                lcs[h] = on[h] ? cb.lc[at] : make_uint2(0, 0);
                oo[h] = on[h] ? (uint32_t) cb.o[at] : 0u;
            }
            uint64_t mm[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                // FB, BB searched the forward index; FF, BF the backward index
                const int ch = firsts[h] ? chA : chB;
                const uint32_t *ri = (ch == 0 || ch == 3) ? fw.reads_index : bw.reads_index;
                mm[h] = on[h] ? entry_read(lcs[h], qq[h] - pre[ee[h]], ri) : 0ull;
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (!on[h])
                    continue;
                const int ch = firsts[h] ? chA : chB;
                item_fill(G, fw, bw, ch, firsts[h] ? 0 : 1, oo[h], mm[h], kw,
                          keys + (uint64_t) qq[h] * kw, meta + qq[h]);
            }
        }
        wave_sync();
        // sorted evaluation: one read per wave sorts to the next power of two of
        // its item count; groups sharing a wave sort to the tier's limit
        uint32_t P = 1;
        while (P < (GW == 64 ? k : kmax))
            P <<= 1;
        extract_sorted<GW>(r, side, kw, k, P, keys, meta, emitf, pre, lcpv, out);
    }
    return true;
}

// Per-group storage: keys (kw words per item), metadata, emit flags, and the
// prefix of entry counts (an entry holds at least one item, so at most kmax);
// with `ent` also both sides' candidate entries (kmax per side), read once from
// the candidate buffers.
__host__ __device__ static inline size_t ext_wave_bytes(uint32_t kmax, uint32_t kw, bool ent = false)
{
    size_t b = (size_t) kmax * kw * 8 + (size_t) kmax * sizeof(ItemMeta) +
               (size_t) (kmax + 2) * 4 + (size_t) kmax * 2 + kmax;
    if (ent)
        b += (size_t) kmax * 2 * (sizeof(uint2) + sizeof(uint16_t));
    return (b + 15) & ~(size_t) 15;
}

struct ExtWave {
    uint64_t *keys;
    ItemMeta *meta;
    uint32_t *pre;
    uint16_t *lcp;       // one-wave tiers: codes shared with the previous sorted item
    uint8_t *emitf;
    uint2 *elc;          // entries of side s at [s * kmax, s * kmax + kmax) (with `ent`)
    uint16_t *eo;
};

__device__ __forceinline__ ExtWave ext_wave(char *base, uint32_t kmax, uint32_t kw,
                                            bool ent = false)
{
    ExtWave w;
    w.keys = (uint64_t *) base;
    char *p = base + (size_t) kmax * kw * 8;
    w.elc = nullptr;
    w.eo = nullptr;
    if (ent) {
        w.elc = (uint2 *) p;
        p += (size_t) kmax * 2 * sizeof(uint2);
    }
    w.meta = (ItemMeta *) p;
    w.pre = (uint32_t *) ((char *) w.meta + (size_t) kmax * sizeof(ItemMeta));
    w.lcp = (uint16_t *) (w.pre + kmax + 2);
    if (ent) {
        w.eo = w.lcp + kmax;
        w.emitf = (uint8_t *) (w.eo + 2 * kmax);
    } else {
        w.emitf = (uint8_t *) (w.lcp + kmax);
    }
    return w;
}

// Process one read with one group of GW lanes whose candidate entries are read
// once into LDS (w.elc / w.eo): both sides' entries in one round of loads, then
// per side only the two dependent loads of its items (reads_index, the key's
// text).  Same results and limits as extract_read.
template <uint32_t KMAX, int GW, typename R>
__device__ bool extract_read_ent(const Geom &G, const DirViewT<R> &fw, const DirViewT<R> &bw,
                                 const ChainBuf &cb, uint64_t rl, uint64_t r, uint32_t kw,
                                 const ExtWave &w, const ExtOut &out, uint32_t &items)
{
    constexpr uint32_t kmax = KMAX;
    constexpr int IPL = (int) (KMAX / GW);   // items (and entries) per lane and side
    const int lane = threadIdx.x & (GW - 1);
    const uint4 c4 = *(const uint4 *) (cb.n + rl * 4);
    const uint32_t cnt4[4] = {c4.x, c4.y, c4.z, c4.w};
    items = 0;
    // side 0 = B (chains 2: BF, 3: BB), side 1 = F (chains 0: FB, 1: FF); an entry
    // holds at least one item, so a side of more than kmax entries overflows
    const uint32_t ne0 = cnt4[2] + cnt4[3], ne1 = cnt4[0] + cnt4[1];
    if (ne0 > kmax || ne1 > kmax)
        return false;
    // both sides' entries in one round of independent loads, then into LDS
    uint32_t k0 = 0, k1 = 0;
    uint2 lcs[2 * IPL];
    uint16_t os[2 * IPL];
#pragma unroll
    for (int h = 0; h < 2 * IPL; h++) {
        const uint32_t e = (uint32_t) lane + (uint32_t) h * GW;
        lcs[h] = make_uint2(0, 0);
        os[h] = 0;
        if (e < ne0 + ne1) {
            const bool s1 = e >= ne0;
            const uint32_t es = s1 ? e - ne0 : e;
            const int chA = s1 ? 0 : 2, chB = s1 ? 1 : 3;
            const uint32_t nA = cnt4[chA];
            const uint64_t at = es < nA ? (rl * 4 + chA) * cb.cap + es
                                        : (rl * 4 + chB) * cb.cap + (es - nA);
            lcs[h] = cb.lc[at];
            os[h] = cb.o[at];
        }
    }
#pragma unroll
    for (int h = 0; h < 2 * IPL; h++) {
        const uint32_t e = (uint32_t) lane + (uint32_t) h * GW;
        if (e < ne0 + ne1) {
            const bool s1 = e >= ne0;
            const uint32_t at = s1 ? kmax + (e - ne0) : e;
            w.elc[at] = lcs[h];
            w.eo[at] = os[h];
            if (s1)
                k1 += entry_count(lcs[h]);
            else
                k0 += entry_count(lcs[h]);
        }
    }
    for (int d = GW / 2; d > 0; d >>= 1) {
        k0 += __shfl_xor(k0, d, GW);
        k1 += __shfl_xor(k1, d, GW);
    }
    if (k0 > kmax || k1 > kmax)
        return false;
    wave_sync();
    for (int side = 0; side < 2; side++) {
        const int chA = side ? 0 : 2, chB = side ? 1 : 3;
        const uint32_t nA = cnt4[chA], ne = side ? ne1 : ne0;
        const uint2 *elc = w.elc + (side ? kmax : 0);
        const uint16_t *eo = w.eo + (side ? kmax : 0);
        // exclusive prefix of the counts over the side's entries
        uint32_t carry = 0;
        for (uint32_t e0 = 0; e0 < ne; e0 += GW) {
            const uint32_t e = e0 + lane;
            const uint32_t v = e < ne ? entry_count(elc[e]) : 0u;
            uint32_t inc = v;
            for (int d = 1; d < GW; d <<= 1) {
                uint32_t x = __shfl_up(inc, d, GW);
                if (lane >= d)
                    inc += x;
            }
            if (e < ne)
                w.pre[e] = carry + inc - v;
            carry += __shfl(inc, GW - 1, GW);
        }
        const uint32_t k = carry;
        items += k;
        wave_sync();
        auto entry_of = [&](uint32_t q) -> uint32_t {   // e with pre[e] <= q < pre[e+1]
            uint32_t lo = 0, hi = ne;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (w.pre[mid] <= q)
                    lo = mid;
                else
                    hi = mid;
            }
            return lo;
        };
        // every item of the lane at once (at most kmax / GW): reads_index, then the
        // key's text, each one round of independent loads
        for (uint32_t q0 = lane; q0 < k; q0 += IPL * GW) {
            uint64_t mm[IPL];
            uint32_t oo[IPL];
            bool firsts[IPL];
#pragma unroll
            for (int h = 0; h < IPL; h++) {
                const uint32_t q = q0 + (uint32_t) h * GW;
                mm[h] = 0;
                oo[h] = 0;
                firsts[h] = true;
                if (q < k) {
                    const uint32_t e = entry_of(q);
                    firsts[h] = e < nA;
                    const uint2 lc = elc[e];
                    oo[h] = eo[e];
                    // FB, BB searched the forward index; FF, BF the backward index
                    const int ch = firsts[h] ? chA : chB;
                    const uint32_t *ri = (ch == 0 || ch == 3) ? fw.reads_index : bw.reads_index;
                    mm[h] = entry_read(lc, q - w.pre[e], ri);
                }
            }
#pragma unroll
            for (int h = 0; h < IPL; h++) {
                const uint32_t q = q0 + (uint32_t) h * GW;
                if (q < k) {
                    const int ch = firsts[h] ? chA : chB;
                    item_fill(G, fw, bw, ch, firsts[h] ? 0 : 1, oo[h], mm[h], kw,
                              w.keys + (uint64_t) q * kw, w.meta + q);
                }
            }
        }
        wave_sync();
        uint32_t P = 1;
        while (P < kmax)
            P <<= 1;
        extract_sorted<GW>(r, side, kw, k, P, w.keys, w.meta, w.emitf, w.pre, w.lcp, out);
    }
    return true;
}

// ---- register tiers ---------------------------------------------------------------------
//
// A read side of at most G * RI items is sorted by a group of G lanes holding
// RI items each in registers (item i at lane i / RI, slot i % RI): the composite
// key (the key's codes left-aligned in KW words, zero past its length; then
// klen << 49 | rank << 48 | o << 32 | m) orders items as item_less does (a key
// that is a proper prefix of another sorts first, which the zero padding and the
// length word give).  The bitonic network compares slots inside a lane and
// exchanges lanes by DPP and ds_swizzle.  Items that are no candidate (contained
// reads, empty slots) carry all-ones keys and sort last.  The emission rules of
// extract_emit then walk the sorted items with every lane of the group holding
// the same state (item values broadcast inside the group).  Small groups keep
// many reads in flight per wave: the extraction waits mostly on its scattered
// reads (reads_index, key text).

#define RI 4

template <int KW> struct RegItem {
    uint64_t w[KW];
    uint64_t tie;
};

template <int KW>
__device__ __forceinline__ bool reg_less(const RegItem<KW> &a, const RegItem<KW> &b)
{
#pragma unroll
    for (int i = 0; i < KW; i++)
        if (a.w[i] != b.w[i])
            return a.w[i] < b.w[i];
    return a.tie < b.tie;
}

template <int J, int KW> __device__ __forceinline__ RegItem<KW> reg_xor(const RegItem<KW> &s)
{
    RegItem<KW> o;
#pragma unroll
    for (int i = 0; i < KW; i++)
        o.w[i] = xor_lane64<J>(s.w[i]);
    o.tie = xor_lane64<J>(s.tie);
    return o;
}

// Step J of merge stage K over the G * RI items of a group.
template <int K, int J, int KW>
__device__ __forceinline__ void reg_step(RegItem<KW> (&it)[RI], int li)
{
    if constexpr (J < RI) {
        // partners inside the lane
#pragma unroll
        for (int h = 0; h < RI; h++) {
            if ((h & J) == 0) {
                const int i = li * RI + h;
                const bool up = (i & K) == 0;
                if (reg_less<KW>(it[h | J], it[h]) == up) {
                    const RegItem<KW> x = it[h];
                    it[h] = it[h | J];
                    it[h | J] = x;
                }
            }
        }
    } else {
        // partner lane li ^ (J / RI), same slot
#pragma unroll
        for (int h = 0; h < RI; h++) {
            const RegItem<KW> p = reg_xor<J / RI, KW>(it[h]);
            const int i = li * RI + h;
            const bool up = (i & K) == 0, lower = (i & J) == 0;
            const bool swap = (lower == up) ? reg_less<KW>(p, it[h]) : reg_less<KW>(it[h], p);
            if (swap)
                it[h] = p;
        }
    }
    if constexpr (J > 1)
        reg_step<K, J / 2, KW>(it, li);
}

template <int N, int K, int KW>
__device__ __forceinline__ void reg_sort(RegItem<KW> (&it)[RI], int li)
{
    reg_step<K, K / 2, KW>(it, li);
    if constexpr (K < N)
        reg_sort<N, 2 * K, KW>(it, li);
}

// The value of lane j of this lane's group of G lanes.
template <int G, int J> __device__ __forceinline__ uint32_t group_bcast(uint32_t v)
{
    if constexpr (G == 4)
        return quad_bcast<J>(v);
    else if constexpr (G <= 32)
        return (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, (J << 5) | (0x1F & ~(G - 1)));
    else
        return (uint32_t) __builtin_amdgcn_readlane((int) v, J);   // the group is the wave
}

// The value of lane li - 1 of the group (lane 0: its own).
template <int G> __device__ __forceinline__ uint64_t group_prev64(uint64_t v)
{
    if constexpr (G == 4) {
        // quad permutation (0, 0, 1, 2)
        const uint32_t lo = (uint32_t) __builtin_amdgcn_mov_dpp((int) (uint32_t) v, 0x90, 0xF, 0xF, false);
        const uint32_t hi = (uint32_t) __builtin_amdgcn_mov_dpp((int) (uint32_t) (v >> 32), 0x90, 0xF, 0xF, false);
        return ((uint64_t) hi << 32) | lo;
    } else {
        return __shfl_up(v, 1, G);
    }
}

// Per-group LDS of the register tiers: both sides' entries and entry prefixes.
template <int N> struct RegEnt {
    uint32_t pre[2][N + 1];
    uint2 lc[2][N];
    uint16_t o[2][N];
};

// State of the ordered pass (extract_emit's rules), the same in every lane of
// a group.
struct EmitState {
    uint32_t M, prevlen, bf, emitted;
};

template <int G, int T, int KW>
__device__ __forceinline__ void reg_emit_steps(const uint32_t (&p1)[RI], const uint32_t (&po)[RI],
                                               uint32_t nv, uint32_t tmax, int li, EmitState &st,
                                               uint32_t (&pos)[RI], uint32_t &keepbits)
{
    constexpr int N = G * RI;
    if (T < (int) tmax) {
        const uint32_t x1 = group_bcast<G, T / RI>(p1[T % RI]);
        const uint32_t x2 = group_bcast<G, T / RI>(po[T % RI]);
        if ((uint32_t) T < nv) {
            const uint32_t Lt = x1 & 511u, At = (x1 >> 9) & 511u, Rt = (x1 >> 18) & 1u,
                           St = (x1 >> 19) & 1u;
            const bool first = T == 0;
            const bool chain = !first && st.M <= At;
            const bool dropped = chain && st.M < Lt;
            const bool same_key = !first && At == Lt && At == st.prevlen;
            const uint32_t ro = (Rt << 16) | x2;
            if (!same_key)
                st.bf = ro;
            const bool keep = !dropped && ro == st.bf && !St;
            st.M = chain ? (st.M < Lt ? st.M : Lt) : Lt;
            st.prevlen = Lt;
            if (keep) {
                if (li == T / RI) {
                    keepbits |= 1u << (T % RI);
                    pos[T % RI] = st.emitted;
                }
                st.emitted++;
            }
        }
        if constexpr (T + 1 < N)
            reg_emit_steps<G, T + 1, KW>(p1, po, nv, tmax, li, st, pos, keepbits);
    }
}

// One read side (side 0 = B: chains BF, BB; side 1 = F: FB, FF) by a group of G
// lanes: ne entries with k items (both <= G * RI) staged in e.  Sorts, applies
// the emission rules, writes the side's counts and edges (a group without
// `write` takes part in the group-wide steps with no items).
template <int G, int KW, typename R>
__device__ __forceinline__ void reg_side(const Geom &Gm, const DirViewT<R> &fw, const DirViewT<R> &bw,
                         uint64_t r, int side, uint32_t nA, uint32_t ne, uint32_t k,
                         const RegEnt<G * RI> &e, bool write, const ExtOut &out)
{
    constexpr int N = G * RI;
    const int lane = (int) (threadIdx.x & 63), li = lane & (G - 1);
    const int chA = side ? 0 : 2, chB = side ? 1 : 3;
    const uint32_t *pre = e.pre[side];
    RegItem<KW> it[RI];
#pragma unroll
    for (int h = 0; h < RI; h++) {
#pragma unroll
        for (int i = 0; i < KW; i++)
            it[h].w[i] = ~0ull;
        it[h].tie = ~0ull;
    }
    // items li * RI + h: entry (the last e with pre[e] <= q), read, key.  Every
    // load is issued unconditionally at a safe address (index 0 or a clamped text
    // word for slots with no item), so that a lane's RI chains of loads are in
    // flight together: a load under a per-slot branch is waited for before the
    // next slot's is issued.
    uint64_t mm[RI];
    uint32_t oo[RI];
    bool fst[RI], act[RI];
    uint2 lch[RI];
    uint64_t ia[RI];
#pragma unroll
    for (int h = 0; h < RI; h++) {
        const uint32_t q = (uint32_t) (li * RI + h);
        act[h] = q < k;
        uint32_t lo = 0, hi = ne;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= q)
                lo = mid;
            else
                hi = mid;
        }
        lch[h] = e.lc[side][lo];
        oo[h] = act[h] ? (uint32_t) e.o[side][lo] : 0u;
        fst[h] = lo < nA;
        ia[h] = act[h] && !(lch[h].y & ENTRY_READ) ? (uint64_t) lch[h].x + (q - pre[lo]) : 0ull;
    }
    uint32_t rv[RI];
#pragma unroll
    for (int h = 0; h < RI; h++) {
        // FB, BB searched the forward index; FF, BF the backward index
        const int ch = fst[h] ? chA : chB;
        rv[h] = ((ch == 0 || ch == 3) ? fw.reads_index : bw.reads_index)[ia[h]];
    }
#pragma unroll
    for (int h = 0; h < RI; h++)
        mm[h] = !act[h] ? 0ull : (lch[h].y & ENTRY_READ) ? (uint64_t) lch[h].x : (uint64_t) rv[h];
    // key m[o..) (FB, BB) or rc(m)[o..) (FF, BF: the complement of the backward
    // text), zero past its length: KW + 1 words per slot
    const uint64_t wmax = (Gm.N >> 5) + 3;   // last slack word of a packed text
    uint64_t raw[RI][KW + 1];
    uint32_t klen[RI], sh[RI];
    bool val[RI], pl[RI];
#pragma unroll
    for (int h = 0; h < RI; h++) {
        const int ch = fst[h] ? chA : chB;
        pl[h] = (ch == 0 || ch == 3);
        const uint32_t Lm = pq_read_len(Gm, mm[h]);
        val[h] = act[h] && oo[h] < Lm;
        klen[h] = val[h] ? Lm - oo[h] : 0u;
        const uint64_t ks = val[h] ? pq_read_start(Gm, mm[h]) + oo[h] : 0ull;
        const uint64_t *tx = pl[h] ? fw.text2 : bw.text2;
        const uint64_t w0 = ks >> 5;
        sh[h] = (uint32_t) (ks & 31) * 2;
#pragma unroll
        for (int i = 0; i <= KW; i++)
            raw[h][i] = tx[w0 + i < wmax ? w0 + i : wmax];
    }
#pragma unroll
    for (int h = 0; h < RI; h++) {
        if (val[h]) {
#pragma unroll
            for (int i = 0; i < KW; i++) {
                uint64_t x = (raw[h][i] << sh[h]) | ((raw[h][i + 1] >> 1) >> (63 - sh[h]));
                if (!pl[h])
                    x = ~x;
                const uint32_t lo = 32u * (uint32_t) i;
                if (lo >= klen[h])
                    x = 0;
                else if (klen[h] - lo < 32)
                    x &= ~0ull << (64 - 2 * (klen[h] - lo));
                it[h].w[i] = x;
            }
            it[h].tie = ((uint64_t) klen[h] << 49) | ((uint64_t) (fst[h] ? 0u : 1u) << 48) |
                        ((uint64_t) oo[h] << 32) | (uint64_t) (uint32_t) mm[h];
        }
    }
    reg_sort<N, 2, KW>(it, li);
    // sorted position li * RI + h: length, codes shared with the previous item
    uint32_t p1[RI], po[RI];
    uint32_t nvl = 0;
#pragma unroll
    for (int h = 0; h < RI; h++) {
        RegItem<KW> pv;
        if (h > 0) {
            pv = it[h - 1];
        } else {
#pragma unroll
            for (int i = 0; i < KW; i++)
                pv.w[i] = group_prev64<G>(it[RI - 1].w[i]);
            pv.tie = group_prev64<G>(it[RI - 1].tie);
        }
        const bool valid = it[h].tie != ~0ull;
        nvl += valid ? 1u : 0u;
        const uint32_t L = valid ? (uint32_t) (it[h].tie >> 49) : 0u;
        const uint32_t Lp = (uint32_t) (pv.tie >> 49);
        const uint32_t mn = L < Lp ? L : Lp;
        uint32_t l = mn;
#pragma unroll
        for (int i = KW - 1; i >= 0; i--) {
            const uint64_t x = pv.w[i] ^ it[h].w[i];
            if (x)
                l = 32u * i + (uint32_t) __clzll((long long) x) / 2;
        }
        const uint32_t A = (li == 0 && h == 0) ? 0u : (l < mn ? l : mn);
        const uint32_t mi = (uint32_t) it[h].tie, rk = (uint32_t) (it[h].tie >> 48) & 1u;
        const uint32_t self = valid && mi == (uint32_t) r ? 1u : 0u;
        p1[h] = L | (A << 9) | (rk << 18) | (self << 19);
        po[h] = (uint32_t) (it[h].tie >> 32) & 0xFFFFu;
    }
    // valid items of the group (they sort first) and the wave's largest count
    uint32_t nv = nvl;
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1)
        nv += __shfl_xor(nv, d, G);
    uint32_t tmax = nv;
#pragma unroll
    for (int d = 32; d >= G; d >>= 1) {
        const uint32_t x = __shfl_xor(tmax, d, 64);
        tmax = x > tmax ? x : tmax;
    }
    tmax = (uint32_t) __builtin_amdgcn_readfirstlane((int) tmax);
    EmitState st = {0, 0, 0, 0};
    uint32_t pos[RI] = {0, 0, 0, 0}, keepbits = 0;
    reg_emit_steps<G, 0, KW>(p1, po, nv, tmax, li, st, pos, keepbits);
    const uint32_t em = st.emitted;
    uint64_t base = 0;
    uint64_t *dst = out.staged + (r * 2 + side) * EDGE_SLOT;
    if (li == 0 && write) {
        if (em > EDGE_SLOT) {
            base = atomicAdd(out.edge_count, (unsigned long long) em);
            if (base + em > out.edge_cap)
                atomicOr(out.flags, 1u);
        }
        if (side) {
            out.ovF[r] = (int32_t) nv;
            out.numF[r] = em;
            out.recF[r] = base;
        } else {
            out.ovB[r] = (int32_t) nv;
            out.numB[r] = em;
            out.recB[r] = base;
        }
    }
    base = ((uint64_t) group_bcast<G, 0>((uint32_t) (base >> 32)) << 32) |
           group_bcast<G, 0>((uint32_t) base);
    uint64_t cap = EDGE_SLOT;
    if (em > EDGE_SLOT) {
        dst = out.edges + base;
        cap = base < out.edge_cap ? out.edge_cap - base : 0;
    }
    if (write) {
#pragma unroll
        for (int h = 0; h < RI; h++) {
            if ((keepbits >> h) & 1u && pos[h] < cap) {
                const uint32_t rk = (uint32_t) (it[h].tie >> 48) & 1u;
                // edge type: FB 1, FF 2 (side F), BF 0, BB 3 (side B)
                const uint32_t type = side ? (rk ? 2u : 1u) : (rk ? 3u : 0u);
                dst[pos[h]] = ((uint64_t) (uint32_t) it[h].tie << 32) |
                              ((uint64_t) po[h] << 8) | type;
            }
        }
    }
}

// One group of G lanes per read (G * RI = 16, 32 or 64 items per side), items
// in registers, 256-thread blocks; reads [0, nr) of the batch or the *count reads
// of `list` (grid-stride, the loop bound the wave's).  A read with a side of
// more than G * RI items goes to out.overflow_reads for the next tier.
// (launch bounds: the 16-item tier, which takes most reads, asks for 4 waves per
// SIMD -- 128 vector registers; the others would spill there)
template <int G, int KW, typename R>
__global__ void __launch_bounds__(256, (G == 4 ? 4 : 1)) k_extract_reg(Geom Gm, DirViewT<R> fw,
                                                                        DirViewT<R> bw,
                                                     ChainBuf cb, uint64_t r0, uint64_t nr,
                                                     const uint32_t *list, const uint32_t *count,
                                                     ExtOut out)
{
    constexpr int N = G * RI, GROUPS = 256 / G;
    __shared__ RegEnt<N> shs[GROUPS];
    const int lane = (int) (threadIdx.x & 63), li = lane & (G - 1);
    const int grp = (int) threadIdx.x / G;
    RegEnt<N> &E = shs[grp];
    const uint64_t total = list ? (uint64_t) *count : nr;
    uint64_t nitems = 0, nreads = 0;
    const uint64_t wave_first =
        (uint64_t) blockIdx.x * GROUPS + (uint64_t) (threadIdx.x / 64) * (64 / G);
    for (uint64_t i0 = wave_first; i0 < total; i0 += (uint64_t) gridDim.x * GROUPS) {
        const uint64_t i = i0 + (uint64_t) (lane / G);
        const bool have = i < total;
        const uint64_t rl = have ? (list ? list[i] : i) : 0;
        uint4 c4 = make_uint4(0, 0, 0, 0);
        if (have)
            c4 = *(const uint4 *) (cb.n + rl * 4);
        // side 0 = B (chains 2, 3), side 1 = F (chains 0, 1); an entry holds at
        // least one item
        const uint32_t ne[2] = {c4.z + c4.w, c4.x + c4.y};
        const uint32_t nA[2] = {c4.z, c4.x};
        bool ok = have && ne[0] <= (uint32_t) N && ne[1] <= (uint32_t) N;
        // both sides' entries (lane li: entries li * RI .. li * RI + RI - 1 of each
        // side) in one round of loads
        // (loaded unconditionally -- a slot with no entry reads its chain's first
        // record -- so that the loads are in flight together)
        uint2 lcs[2][RI];
        uint16_t os[2][RI];
#pragma unroll
        for (int sd = 0; sd < 2; sd++)
#pragma unroll
            for (int j = 0; j < RI; j++) {
                const uint32_t en = (uint32_t) (li * RI + j);
                const int cA = sd ? 0 : 2, cB = sd ? 1 : 3;
                const uint64_t at = !(ok && en < ne[sd]) ? (rl * 4 + cA) * cb.cap
                                    : en < nA[sd]        ? (rl * 4 + cA) * cb.cap + en
                                                  : (rl * 4 + cB) * cb.cap + (en - nA[sd]);
                lcs[sd][j] = cb.lc[at];
                os[sd][j] = cb.o[at];
            }
#pragma unroll
        for (int sd = 0; sd < 2; sd++)
#pragma unroll
            for (int j = 0; j < RI; j++)
                if (!(ok && (uint32_t) (li * RI + j) < ne[sd])) {
                    lcs[sd][j] = make_uint2(0, 0);
                    os[sd][j] = 0;
                }
        // items per side and each entry's exclusive prefix, into LDS
        uint32_t k[2];
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            uint32_t part = 0;
#pragma unroll
            for (int j = 0; j < RI; j++)
                part += (uint32_t) (li * RI + j) < ne[sd] ? entry_count(lcs[sd][j]) : 0u;
            uint32_t inc = part;
#pragma unroll
            for (int d = 1; d < G; d <<= 1) {
                const uint32_t x = __shfl_up(inc, d, G);
                if (li >= d)
                    inc += x;
            }
            k[sd] = (uint32_t) __shfl((int) inc, G - 1, G);
            uint32_t run = inc - part;
#pragma unroll
            for (int j = 0; j < RI; j++) {
                const uint32_t en = (uint32_t) (li * RI + j);
                if (en < ne[sd]) {
                    E.pre[sd][en] = run;
                    E.lc[sd][en] = lcs[sd][j];
                    E.o[sd][en] = os[sd][j];
                    run += entry_count(lcs[sd][j]);
                }
            }
        }
        ok = ok && k[0] <= (uint32_t) N && k[1] <= (uint32_t) N;
        wave_sync();
        const uint64_t r = r0 + rl;
        reg_side<G, KW, R>(Gm, fw, bw, r, 0, nA[0], ok ? ne[0] : 0u, ok ? k[0] : 0u, E, ok, out);
        reg_side<G, KW, R>(Gm, fw, bw, r, 1, nA[1], ok ? ne[1] : 0u, ok ? k[1] : 0u, E, ok, out);
        wave_sync();   // the next reads overwrite the entries
        if (ok) {
            nitems += k[0] + k[1];
            nreads++;
        }
        // reads left for the next tier, appended with one atomic per wave
        const bool ov = have && !ok && li == 0;
        const uint64_t bal = __ballot(ov);
        if (bal) {
            const int lead = __ffsll((long long) bal) - 1;
            uint32_t b = 0;
            if (lane == lead)
                b = atomicAdd(out.n_overflow, (uint32_t) __popcll(bal));
            b = __shfl(b, lead, 64);
            if (ov)
                out.overflow_reads[b + (uint32_t) __popcll(bal & ((1ull << lane) - 1))] =
                    (uint32_t) rl;
        }
    }
    if (out.xstats)
        ext_count<G>(out, nitems, nreads);
}

// One group of GW lanes per read, storage in LDS, WAVES waves per block.  Reads
// [0, nr) of the batch, or (list != nullptr) the *count reads of `list`, taken
// grid-stride so that the grid need not match the count known only on the
// device.  A read with more than KMAX items on a side is appended to
// out.overflow_reads for the next tier.
template <uint32_t KMAX, int WAVES, int GW, typename R>
__global__ void __launch_bounds__(64 * WAVES) k_extract_lds(Geom G, DirViewT<R> fw,
                                                           DirViewT<R> bw, ChainBuf cb,
                                                           uint64_t r0, uint64_t nr,
                                                           const uint32_t *list,
                                                           const uint32_t *count, uint32_t kw,
                                                           ExtOut out)
{
    extern __shared__ uint64_t lds[];
    constexpr int GROUPS = WAVES * 64 / GW;
    const int grp = threadIdx.x / GW;
    const uint64_t total = list ? (uint64_t) *count : nr;
    constexpr bool ENT = GW < 64;
    ExtWave w = ext_wave((char *) lds + ext_wave_bytes(KMAX, kw, ENT) * grp, KMAX, kw, ENT);
    uint64_t nitems = 0, nreads = 0;
    for (uint64_t i = (uint64_t) blockIdx.x * GROUPS + grp; i < total;
         i += (uint64_t) gridDim.x * GROUPS) {
        uint64_t rl = list ? list[i] : i;
        uint32_t items;
        bool ok;
        if constexpr (ENT)
            ok = extract_read_ent<KMAX, GW, R>(G, fw, bw, cb, rl, r0 + rl, kw, w, out, items);
        else
            ok = extract_read<GW, R>(G, fw, bw, cb, rl, r0 + rl, kw, w.keys, w.meta, w.emitf,
                                     w.pre, w.lcp, KMAX, out, items);
        if (ok) {
            nitems += items;
            nreads++;
        }
        // reads left for the next tier, appended with one atomic per wave
        const bool ov = !ok && (threadIdx.x & (GW - 1)) == 0;
        const uint64_t bal = __ballot(ov);
        if (bal) {
            const int lane = threadIdx.x & 63, lead = __ffsll((long long) bal) - 1;
            uint32_t base = 0;
            if (lane == lead)
                base = atomicAdd(out.n_overflow, (uint32_t) __popcll(bal));
            base = __shfl(base, lead, 64);
            if (ov)
                out.overflow_reads[base + (uint32_t) __popcll(bal & ((1ull << lane) - 1))] =
                    (uint32_t) rl;
        }
    }
    if (out.xstats)
        ext_count<GW>(out, nitems, nreads);
}

// ---- read sides beyond the global tier ------------------------------------------------
//
// A side with more than KMAX_GLOBAL items (a read inside a high-copy repeat) is
// listed by the global tier with the item counts of both sides.  Eager
// searches then sort each side of such a read with all of the GPU: the items
// are expanded into scratch sized to the side, sorted by a bitonic network of
// one launch per stage, and the emission rules run over the sorted order by
// one wave (extract_emit, as every tier).  There is no limit on candidates but
// device memory; the reference keeps any number of candidates as SA intervals
// (string_graph.c:773-778).

struct HugeRead {
    uint32_t rl, pad;
    uint64_t k[2];       // items of side 0 (B) and side 1 (F)
};

// Items of one side of batch read rl, summed by one group of GW lanes.
template <int GW> __device__ uint64_t side_items(const ChainBuf &cb, uint64_t rl, int side)
{
    const int lane = threadIdx.x & (GW - 1);
    const int chA = side ? 0 : 2, chB = side ? 1 : 3;
    const uint32_t nA = cb.n[rl * 4 + chA], nB = cb.n[rl * 4 + chB];
    uint64_t k = 0;
    for (uint32_t e = lane; e < nA + nB; e += GW) {
        const uint64_t at = e < nA ? (rl * 4 + chA) * cb.cap + e : (rl * 4 + chB) * cb.cap + (e - nA);
        k += entry_count(cb.lc[at]);
    }
    for (int d = GW / 2; d > 0; d >>= 1)
        k += __shfl_xor(k, d, GW);
    return k;
}

// One wave per read with scratch in global memory (up to KMAX_GLOBAL items per
// side).  A larger read is listed in `huge` (eager searches), or, while the
// search is being captured as a graph (huge_is_error), raises flag 2: the host
// then runs the search eagerly.
template <typename R>
__global__ void __launch_bounds__(64) k_extract_global(Geom G, DirViewT<R> fw, DirViewT<R> bw,
                                                       ChainBuf cb, uint64_t r0,
                                                       const uint32_t *list,
                                                       const uint32_t *count, uint32_t kw,
                                                       char *gscratch, ExtOut out, HugeRead *huge,
                                                       uint32_t *n_huge, uint32_t huge_is_error)
{
    ExtWave w = ext_wave(gscratch + ext_wave_bytes(KMAX_GLOBAL, kw) * blockIdx.x, KMAX_GLOBAL, kw);
    const uint32_t nlist = *count;
    uint64_t nitems = 0, nreads = 0;
    for (uint32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
        uint64_t rl = list[i];
        uint32_t items;
        bool ok = extract_read<64, R>(G, fw, bw, cb, rl, r0 + rl, kw, w.keys, w.meta, w.emitf,
                                      w.pre, w.lcp, KMAX_GLOBAL, out, items);
        if (ok) {
            nitems += items;
            nreads++;
            continue;
        }
        const uint64_t kB = side_items<64>(cb, rl, 0), kF = side_items<64>(cb, rl, 1);
        if (threadIdx.x == 0) {
            if (huge_is_error) {
                atomicOr(out.flags, 2u);
            } else {
                const uint32_t at = atomicAdd(n_huge, 1u);
                HugeRead h;
                h.rl = (uint32_t) rl;
                h.pad = 0;
                h.k[0] = kB;
                h.k[1] = kF;
                huge[at] = h;
            }
        }
    }
    if (out.xstats)
        ext_count<64>(out, nitems, nreads);
}

// Exclusive prefix of the item counts over the entries of one side; pre[ne] =
// the side's items.  One wave.
__global__ void __launch_bounds__(64) k_huge_prefix(ChainBuf cb, uint64_t rl, int side, uint32_t *pre)
{
    const int lane = threadIdx.x;
    const int chA = side ? 0 : 2, chB = side ? 1 : 3;
    const uint32_t nA = cb.n[rl * 4 + chA], ne = nA + cb.n[rl * 4 + chB];
    uint32_t carry = 0;
    for (uint32_t e0 = 0; e0 < ne; e0 += 64) {
        const uint32_t e = e0 + lane;
        uint32_t v = 0;
        if (e < ne) {
            const uint64_t at = e < nA ? (rl * 4 + chA) * cb.cap + e : (rl * 4 + chB) * cb.cap + (e - nA);
            v = entry_count(cb.lc[at]);
        }
        uint32_t inc = v;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t x = __shfl_up(inc, d, 64);
            if (lane >= d)
                inc += x;
        }
        if (e < ne)
            pre[e] = carry + inc - v;
        carry += __shfl(inc, 63, 64);
    }
    if (lane == 0)
        pre[ne] = carry;
}

// Items q < k of one side with their keys; perm[q] = q for all q < P.
template <typename R>
__global__ void k_huge_expand(Geom G, DirViewT<R> fw, DirViewT<R> bw, ChainBuf cb, uint64_t rl,
                              int side, const uint32_t *pre, uint32_t k, uint32_t P, uint32_t kw,
                              uint64_t *keys, ItemMeta *meta, uint32_t *perm)
{
    const int chA = side ? 0 : 2, chB = side ? 1 : 3;
    const uint32_t nA = cb.n[rl * 4 + chA], ne = nA + cb.n[rl * 4 + chB];
    PQ_GRID_STRIDE_FOR(q, P)
    {
        perm[q] = (uint32_t) q;
        if (q >= k)
            continue;
        uint32_t lo = 0, hi = ne;   // the entry e with pre[e] <= q < pre[e + 1]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= q)
                lo = mid;
            else
                hi = mid;
        }
        const bool first = lo < nA;
        const int ch = first ? chA : chB;
        const uint64_t at = first ? (rl * 4 + chA) * cb.cap + lo : (rl * 4 + chB) * cb.cap + (lo - nA);
        const uint2 lc = cb.lc[at];
        const uint32_t *ri = (ch == 0 || ch == 3) ? fw.reads_index : bw.reads_index;
        const uint64_t m = entry_read(lc, (uint32_t) q - pre[lo], ri);
        item_fill(G, fw, bw, ch, first ? 0u : 1u, (uint32_t) cb.o[at], m, kw,
                  keys + (uint64_t) q * kw, meta + q);
    }
}

// One stage (s2, j2) of the bitonic network over perm[0..P).
__global__ void k_huge_sort_step(uint32_t P, uint32_t s2, uint32_t j2, uint32_t k, uint32_t kw,
                                 const uint64_t *keys, const ItemMeta *meta, uint32_t *perm)
{
    PQ_GRID_STRIDE_FOR(p, P / 2)
    {
        const uint32_t pp = (uint32_t) p;
        const uint32_t i = 2 * j2 * (pp / j2) + (pp % j2), ip = i + j2;
        const bool up = (i & s2) == 0;
        const uint32_t x = perm[i], y = perm[ip];
        if (item_less(y, x, k, kw, keys, meta) == up) {
            perm[i] = y;
            perm[ip] = x;
        }
    }
}

__global__ void __launch_bounds__(64) k_huge_emit(uint64_t r, int side, uint32_t kw, uint32_t k,
                                                  uint32_t P, const uint64_t *keys,
                                                  const ItemMeta *meta, uint8_t *emitf,
                                                  const uint32_t *perm, uint16_t *lcpv, ExtOut out)
{
    extract_emit<64>(r, side, kw, k, P, keys, meta, emitf, perm, lcpv, out);
}

// Zero `count` words (a kernel rather than a memset node, so that the sequence
// stays capturable).
__global__ void k_zero_u32(uint32_t *p, uint32_t count)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] = 0;
}

// tot[0..n) = edges per read, tot[n] = 0 so that the exclusive scan of n+1
// entries ends with the total
__global__ void k_edge_totals(const uint32_t *nb, const uint32_t *nf, uint64_t n, uint64_t *tot)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        tot[i] = (uint64_t) nb[i] + nf[i];
    else if (i == n)
        tot[i] = 0;
}

// Nothing is copied when the search raised an error flag (edge buffer
// exhausted: the counts then exceed what the buffers hold; candidate overflow):
// the result is discarded, and the copy would read and write past the buffers.
__global__ void k_csr(const uint64_t *recB, const uint64_t *recF, const uint32_t *numB,
                      const uint32_t *numF, const uint64_t *off, const uint64_t *staged,
                      const uint64_t *pool, uint64_t n, const uint32_t *flags, uint64_t *dst)
{
    uint64_t r = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || (*flags & 3u))
        return;
    uint64_t a = off[r], nb = numB[r], nf = numF[r];
    const uint64_t *sb = nb > EDGE_SLOT ? pool + recB[r] : staged + (r * 2) * EDGE_SLOT;
    const uint64_t *sf = nf > EDGE_SLOT ? pool + recF[r] : staged + (r * 2 + 1) * EDGE_SLOT;
    for (uint64_t i = 0; i < nb; i++)
        dst[a + i] = sb[i];
    for (uint64_t i = 0; i < nf; i++)
        dst[a + nb + i] = sf[i];
}

// Edges in the reference's edge_t layout (string_graph.h: INDEX_TYPE_T dest,
// unsigned short overlap_len, char type, padded to the alignment of dest).
__global__ void k_ref_edges(const uint64_t *pk, uint64_t E, uint32_t index_bytes, uint8_t *out)
{
    uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E)
        return;
    const uint64_t e = pk[i];
    const uint64_t dest = e >> 32;
    const uint16_t len = (uint16_t) ((e >> 8) & 0xFFFF);
    const uint8_t type = (uint8_t) (e & 0xFF);
    if (index_bytes == 4) {
        // 8 bytes: dest u32 | len u16 | type u8 | pad
        uint2 v = make_uint2((uint32_t) dest, (uint32_t) len | ((uint32_t) type << 16));
        ((uint2 *) out)[i] = v;
    } else {
        // 16 bytes: dest u64 | len u16 | type u8 | pad
        uint4 v = make_uint4((uint32_t) dest, (uint32_t) (dest >> 32),
                             (uint32_t) len | ((uint32_t) type << 16), 0u);
        ((uint4 *) out)[i] = v;
    }
}

// Node records, string_graph.c:842-889: SUSPECT mark from the candidate counts
// (ratio in double, as the reference computes it), B/F types from the edge
// counts, REMOVED for a node without edges.  Flags a side of more than 65535
// edges (the 16-bit counts of node_t).
__global__ void k_node_records(const int32_t *ovB, const int32_t *ovF, const uint32_t *nB,
                               const uint32_t *nF, uint64_t n, uint2 *out, uint32_t *over)
{
    PQ_GRID_STRIDE_FOR(i, n)
    {
        const uint32_t b = nB[i], f = nF[i];
        if (b > 65535u || f > 65535u)
            atomicOr(over, 1u);
        const int32_t a = ovB[i], c = ovF[i], mx = a > c ? a : c;
        const double ratio = mx == 0 ? 0.0 : (double) abs(a - c) / (double) mx;
        const uint32_t marked = ratio > 0.6 ? 4u : 0u;
        uint32_t bt = 4u, ft = 4u, status = 8u;          // ENDING, ENDING, REMOVED
        if (b || f) {
            bt = b == 1 ? 1u : b > 1 ? 2u : 4u;
            ft = f == 1 ? 1u : f > 1 ? 2u : 4u;
            status = 0u;
        }
        out[i] = make_uint2((b & 0xFFFFu) | (f << 16), bt | (ft << 8) | (status << 16) | (marked << 24));
    }
}

// Sum of the overlap lengths of all edges (packed dest<<32 | len<<8 | type).
__global__ void k_edge_len_sum(const uint64_t *pk, uint64_t E, unsigned long long *sum)
{
    uint64_t s = 0;
    PQ_GRID_STRIDE_FOR(i, E)
        s += (pk[i] >> 8) & 0xFFFFu;
    for (int o = 32; o; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63u) == 0 && s)
        atomicAdd(sum, (unsigned long long) s);
}

__global__ void k_zero_words(unsigned long long *p, uint32_t count)
{
    if (threadIdx.x < count)
        p[threadIdx.x] = 0ull;
}

// Scan input of one batch's part of the CSR offsets: in[0] = the offset where the
// batch's edges start (final once the batches before it are done), in[1 + j] =
// edges of batch read j, in[nr + 1] = 0.  Entries 1 .. nr + 1 of its exclusive
// scan are edge_off[r0 .. r0 + nr].
__global__ void k_batch_totals(const uint32_t *nb, const uint32_t *nf, uint64_t nr,
                               const uint64_t *base, uint64_t *in)
{
    const uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        in[0] = *base;
    else if (i <= nr)
        in[i] = (uint64_t) nb[i - 1] + nf[i - 1];
    else if (i == nr + 1)
        in[i] = 0;
}

// Internal results of one search attempt (positive: pq_find_overlaps retries).
#define PQ_RETRY_EAGER 1     // a graph replay met a side beyond the global tier
#define PQ_RETRY_EDGES 2     // the shared edge pool overflowed; ctx->edge_need = demand

// The listed reads of one batch whose sides exceed the global tier: each side is
// expanded into scratch sized to it, sorted by one launch per bitonic stage over
// the whole GPU, and its edges emitted by one wave (eager searches only).
template <typename R>
static int extract_huge(pq_ctx_t *ctx, const Geom &G, const DirViewT<R> &fw, const DirViewT<R> &bw,
                        const ChainBuf &cb, uint64_t r0, uint32_t kw, const ExtOut &eo,
                        const HugeRead *list, uint32_t count)
{
    for (uint32_t h = 0; h < count; h++) {
        for (int side = 0; side < 2; side++) {
            const uint64_t k = list[h].k[side];
            if (k >= (1ull << 31))
                return ctx->fail(PQ_ERR_LIMIT, "a read side has 2^31 overlap candidates or more");
            uint32_t P = 1;
            while (P < k)
                P <<= 1;
            auto up16 = [](size_t b) { return (b + 15) & ~(size_t) 15; };
            const size_t b_keys = up16((size_t) P * kw * 8), b_meta = up16((size_t) P * sizeof(ItemMeta));
            const size_t b_perm = up16((size_t) P * 4), b_lcp = up16((size_t) P * 2);
            const size_t b_emit = up16(P), b_pre = up16((size_t) (2 * cb.cap + 2) * 4);
            char *base = nullptr;
            int rc = pq_slot(ctx, SLOT_HUGE, b_keys + b_meta + b_perm + b_lcp + b_emit + b_pre, &base);
            if (rc)
                return rc;
            uint64_t *keys = (uint64_t *) base;
            ItemMeta *meta = (ItemMeta *) (base + b_keys);
            uint32_t *perm = (uint32_t *) (base + b_keys + b_meta);
            uint16_t *lcpv = (uint16_t *) (base + b_keys + b_meta + b_perm);
            uint8_t *emitf = (uint8_t *) (base + b_keys + b_meta + b_perm + b_lcp);
            uint32_t *pre = (uint32_t *) (base + b_keys + b_meta + b_perm + b_lcp + b_emit);
            const uint64_t rl = list[h].rl;
            hipLaunchKernelGGL(k_huge_prefix, dim3(1), dim3(64), 0, ctx->stream, cb, rl, side, pre);
            hipLaunchKernelGGL(k_huge_expand<R>, dim3(pq_blocks_gs(P, 256)), dim3(256), 0,
                               ctx->stream, G, fw, bw, cb, rl, side, pre, (uint32_t) k, P, kw,
                               keys, meta, perm);
            for (uint32_t s2 = 2; s2 <= P; s2 <<= 1)
                for (uint32_t j2 = s2 >> 1; j2 > 0; j2 >>= 1)
                    hipLaunchKernelGGL(k_huge_sort_step, dim3(pq_blocks_gs(P / 2, 256)), dim3(256),
                                       0, ctx->stream, P, s2, j2, (uint32_t) k, kw, keys, meta,
                                       perm);
            hipLaunchKernelGGL(k_huge_emit, dim3(1), dim3(64), 0, ctx->stream, r0 + rl, side, kw,
                               (uint32_t) k, P, keys, meta, emitf, perm, lcpv, eo);
            PQ_CHECK(hipGetLastError());
        }
    }
    return PQ_OK;
}

static double ev_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    (void) hipEventElapsedTime(&ms, a, b);
    return ms;
}

void pq_search_graph_free(pq_ctx_t *ctx)
{
    if (ctx->search_exec)
        (void) hipGraphExecDestroy(ctx->search_exec);
    if (ctx->search_graph)
        (void) hipGraphDestroy(ctx->search_graph);
    ctx->search_exec = nullptr;
    ctx->search_graph = nullptr;
    ctx->search_key.clear();
}

// The search's work counters from its misc words: WSTAT_WORDS 64-bit sums (low
// word first) at misc[8..39], in WS_* order.
static void search_counters(pq_ctx_t *ctx, const uint32_t *hm)
{
    auto u64 = [&](int q) { return (uint64_t) hm[8 + 2 * q] | ((uint64_t) hm[9 + 2 * q] << 32); };
    ctx->timing.find_positions = u64(WS_POSITIONS);
    ctx->timing.find_lines = u64(WS_LINES);
    ctx->timing.verify_rows = u64(WS_VERIFY_ROWS);
    ctx->timing.seed_rows = u64(WS_SEED_ROWS);
    ctx->timing.extract_items = u64(WS_ITEMS);
    for (int k = 0; k < 4; k++)
        ctx->timing.extract_reads[k] = u64(WS_TIER_READS + k);
}

// Graph mode of pq_find_overlaps: the search sequence plus the edge CSR (into an
// edge array sized for the worst case) as one graph; recaptured when any launch
// argument it depends on changes.
template <typename Enqueue>
static int run_search_graph(pq_ctx_t *ctx, int mode, uint64_t n, uint64_t batch, uint64_t edge_cap,
                            uint32_t *misc, uint64_t *rec, uint64_t *recF, uint64_t *staged,
                            uint64_t *app, uint64_t *tot, Enqueue &enqueue,
                            std::initializer_list<uint64_t> args)
{
    int rc;
    const uint64_t ub = 2ull * EDGE_SLOT * n + edge_cap + 1;
    if ((rc = pq_slot(ctx, SLOT_EDGES, ub * 8, &ctx->d_edges)))
        return rc;
    if (!ctx->tmp(SLOT_SCAN, pq_scan_scratch_bytes(n + 1, 8)))
        return ctx->fail(PQ_ERR_NOMEM, "scan scratch");
    std::vector<uint64_t> key = {(uint64_t) mode, ctx->ext_shape, n, batch, edge_cap,
                                 mode == PQ_SEARCH_LEARNED ? ctx->learned_gen : 0,
                                 (uint64_t) (uintptr_t) misc, (uint64_t) (uintptr_t) rec,
                                 (uint64_t) (uintptr_t) recF, (uint64_t) (uintptr_t) staged,
                                 (uint64_t) (uintptr_t) app, (uint64_t) (uintptr_t) tot,
                                 (uint64_t) (uintptr_t) ctx->d_edges,
                                 (uint64_t) (uintptr_t) ctx->d_edge_off,
                                 (uint64_t) (uintptr_t) ctx->d_ovB, (uint64_t) (uintptr_t) ctx->d_ovF,
                                 (uint64_t) (uintptr_t) ctx->d_numB,
                                 (uint64_t) (uintptr_t) ctx->d_numF,
                                 (uint64_t) (uintptr_t) ctx->scratch[SLOT_SCAN].p};
    key.insert(key.end(), args.begin(), args.end());
    for (int d = 0; d < 2; d++) {
        const DirIndex &D = ctx->dir[d];
        key.insert(key.end(), {(uint64_t) (uintptr_t) D.text2, (uint64_t) (uintptr_t) D.sa,
                               (uint64_t) (uintptr_t) D.occ, (uint64_t) (uintptr_t) D.reads_index});
        for (int c = 0; c < 5; c++)
            key.push_back(D.cf[c]);
    }
    if (!ctx->search_exec || key != ctx->search_key) {
        pq_search_graph_free(ctx);
        PQ_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        rc = enqueue(false);
        if (rc == PQ_OK) {
            hipLaunchKernelGGL(k_edge_totals, dim3(pq_blocks(n + 1, 256)), dim3(256), 0,
                               ctx->stream, ctx->d_numB, ctx->d_numF, n, tot);
            rc = pq_scan_u64(ctx, tot, ctx->d_edge_off, n + 1, nullptr);
        }
        if (rc == PQ_OK)
            hipLaunchKernelGGL(k_csr, dim3(pq_blocks(n, 256)), dim3(256), 0, ctx->stream, rec,
                               recF, ctx->d_numB, ctx->d_numF, ctx->d_edge_off, staged, app, n,
                               misc + 3, ctx->d_edges);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc)
            return rc;
        if (e != hipSuccess)
            return ctx->fail(PQ_ERR_HIP, std::string("search capture: ") + hipGetErrorString(e));
        ctx->search_graph = g;
        PQ_CHECK(hipGraphInstantiate(&ctx->search_exec, g, nullptr, nullptr, 0));
        ctx->search_key = key;
    }
    PQ_CHECK(hipGraphLaunch(ctx->search_exec, ctx->stream));
    uint32_t hm[40];
    uint64_t total = 0;
    PQ_CHECK(hipMemcpyAsync(hm, misc, sizeof(hm), hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipMemcpyAsync(&total, ctx->d_edge_off + n, 8, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    search_counters(ctx, hm);
    if (hm[3] & 2u)
        return PQ_RETRY_EAGER;
    if (hm[3] & 1u) {
        ctx->edge_need = (uint64_t) hm[0] | ((uint64_t) hm[1] << 32);
        return PQ_RETRY_EDGES;
    }
    ctx->n_edges = total;
    ctx->searched = true;
    return PQ_OK;
}

// The graph exported into caller memory during an eager search: once a batch's
// edges are extracted, its part of the CSR, its node records and its edge_t
// records are made on the device and leave for the host on the transfer stream
// while the next batch is searched.
struct ExportSpec {
    pq_node_t *nodes;          // n records, or nullptr
    uint8_t *edges;            // edge_capacity edge_t records, or nullptr
    uint64_t edge_capacity;
    uint32_t index_bytes;      // 4 or 8
    uint64_t len_sum;          // out: sum of the overlap lengths
    bool edges_short;          // out: the edges did not fit edge_capacity
    bool reserve_only;         // allocate what the search needs, launch nothing
};

template <typename R>
static int find_overlaps_t(pq_ctx_t *ctx, int mode, bool as_graph, ExportSpec *xp = nullptr)
{
    const uint64_t n = ctx->n;
    const uint32_t t = ctx->t;
    ctx->searched = false;
    ctx->timing.find_ms = ctx->timing.extract_ms = ctx->timing.compact_ms = 0;
    ctx->timing.learned_ms = ctx->timing.seed_kernel_ms = 0;
    ctx->timing.find_kernel_ms = 0;
    ctx->timing.find_launches = 0;
    ctx->timing.find_positions = ctx->timing.find_lines = ctx->timing.verify_rows = 0;
    ctx->timing.seed_rows = ctx->timing.extract_items = 0;
    for (int k = 0; k < 4; k++)
        ctx->timing.extract_reads[k] = 0;
    ctx->n_edges = 0;
    int rc;
    if ((rc = pq_slot(ctx, SLOT_OVB, (n + 1) * 4, &ctx->d_ovB)) ||
        (rc = pq_slot(ctx, SLOT_OVF, (n + 1) * 4, &ctx->d_ovF)) ||
        (rc = pq_slot(ctx, SLOT_NUMB, (n + 1) * 4, &ctx->d_numB)) ||
        (rc = pq_slot(ctx, SLOT_NUMF, (n + 1) * 4, &ctx->d_numF)) ||
        (rc = pq_slot(ctx, SLOT_EDGE_OFF, (n + 1) * 8, &ctx->d_edge_off)))
        return rc;
    if (n == 0) {
        PQ_CHECK(hipMemsetAsync(ctx->d_edge_off, 0, 8, ctx->stream));
        ctx->searched = true;
        return PQ_OK;
    }
    Geom G = ctx->geom();
    const DirViewT<R> fw = ctx->view<R>(0), bw = ctx->view<R>(1);
    const CTab<R> Ct = ctx->ctab<R>();
    ChainBuf cb;
    cb.cap = ctx->Lmax > t ? ctx->Lmax - t : 1;
    uint32_t kw = (cb.cap + 31) / 32;
    if (kw > 8)
        return ctx->fail(PQ_ERR_LIMIT, "extension keys longer than 256 codes");
    // batch so the candidate buffers take at most a third of the free device
    // memory (at least 8 GiB): fewer, larger batches keep the GPU full to the
    // end of each launch (chains of uneven reads finish at very different times)
    uint64_t per_read = 4ull * cb.cap * (sizeof(uint2) + sizeof(uint16_t)) + 16 + 4 * 12;
    uint64_t budget = 8ull << 30;
    {
        size_t free_b = 0, total_b = 0;
        const size_t held = ctx->scratch.size() > SLOT_CAND_O
                                ? ctx->scratch[SLOT_CAND_LC].bytes + ctx->scratch[SLOT_CAND_O].bytes
                                : 0;
        // buffers this search reuses count as free
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            budget = std::max<uint64_t>(budget, ((uint64_t) free_b + held) / 3);
    }
    uint64_t batch = std::max<uint64_t>(1, std::min<uint64_t>(n, budget / per_read));
    if (ctx->search_batch)
        batch = std::min<uint64_t>(batch, ctx->search_batch);
    uint64_t *rec = nullptr, *recF = nullptr, *app = nullptr, *tot = nullptr, *staged = nullptr;
    uint32_t *ovl = nullptr, *misc = nullptr;
    unsigned long long *wstats = nullptr;
    // shared edge pool: the set or default capacity, or what an overflowing
    // search of these reads asked for
    const uint64_t edge_cap = std::max<uint64_t>(
        ctx->edge_cap_set ? ctx->edge_cap_set : std::max<uint64_t>(1 << 20, n / 4), ctx->edge_cap_auto);
    SeedT<R> *seeds = nullptr;
    char *gscratch = nullptr;
    HugeRead *huge = nullptr;
    // five events per batch (seed, find, extract boundaries; the seed kernel's start
    // after the model's prediction), read after the loop
    // batch boundaries: equal batches of at most `batch` reads, then (above a
    // million reads) a batch of about a quarter of the others, so that a search
    // that exports as it goes has little transfer left after its last kernels
    // (every search is batched alike: the same launches with or without export)
    std::vector<uint64_t> bnd(1, 0);
    {
        uint64_t head = n, last = 0;
        if (n >= (1u << 20)) {
            const uint64_t k = (4 * n + 5 * batch - 1) / (5 * batch) > 0
                                   ? (4 * n + 5 * batch - 1) / (5 * batch)
                                   : 1;                       // full batches, last ~ size / 4
            const uint64_t size = (4 * n + 4 * k) / (4 * k + 1);
            last = n - std::min<uint64_t>(n - 1, k * size);
            head = n - last;
        }
        const uint64_t nh = (head + batch - 1) / batch;
        for (uint64_t b = 1; b <= nh; b++)
            bnd.push_back(head * b / nh);
        if (last)
            bnd.push_back(n);
    }
    const uint64_t nbatch = bnd.size() - 1;
    std::vector<hipEvent_t> ev(5 * nbatch);
    for (auto &e : ev)
        (void) hipEventCreate(&e);
    auto cleanup = [&]() {
        for (auto &e : ev)
            (void) hipEventDestroy(e);
    };
    cb.lc = nullptr;
    cb.o = nullptr;
    cb.n = nullptr;
    // verification takes chains of up to vmax rows; their candidates (one entry per
    // read, at most one per row) must fit the entries of a chain
    cb.vmax = std::min<uint32_t>(std::min<uint32_t>(ctx->verify_rows, VERIFY_MAX), cb.cap);
    if ((rc = pq_slot(ctx, SLOT_VJOB_L, batch * 4 * 8, &cb.vl)) ||
        (rc = pq_slot(ctx, SLOT_VJOB, batch * 4 * 4, &cb.vjob)) ||
        (rc = pq_slot(ctx, SLOT_CAND_LC, batch * 4 * cb.cap * sizeof(uint2), &cb.lc)) ||
        (rc = pq_slot(ctx, SLOT_CAND_O, batch * 4 * cb.cap * sizeof(uint16_t), &cb.o)) ||
        (rc = pq_slot(ctx, SLOT_CAND_N, batch * 4 * sizeof(uint32_t), &cb.n)) ||
        (rc = pq_slot(ctx, SLOT_RECB, (n + 1) * 8, &rec)) ||
        (rc = pq_slot(ctx, SLOT_RECF, (n + 1) * 8, &recF)) ||
        (rc = pq_slot(ctx, SLOT_APPEND, edge_cap * 8, &app)) ||
        (rc = pq_slot(ctx, SLOT_STAGED, (n + 1) * 2 * EDGE_SLOT * 8, &staged)) ||
        (rc = pq_slot(ctx, SLOT_TOT, (n + 1) * 8, &tot)) ||
        (rc = pq_slot(ctx, SLOT_OVL, (5 * batch + 5) * 4, &ovl)) ||
        (rc = pq_slot(ctx, SLOT_MISC, 192, &misc)) ||
        (rc = pq_slot(ctx, SLOT_WSTATS, (size_t) WSTAT_SLOTS * WSTAT_STRIDE * 8, &wstats)) ||
        (rc = pq_slot(ctx, SLOT_GSCRATCH, GLOBAL_SLOTS * ext_wave_bytes(KMAX_GLOBAL, kw),
                      &gscratch)) ||
        (rc = pq_slot(ctx, SLOT_HUGE_LIST, batch * sizeof(HugeRead), &huge)) ||
        (rc = pq_slot(ctx, SLOT_WORKQ,
                      ((mode == PQ_SEARCH_LEARNED ? batch * 4 : 0) + 64) * 4, &cb.wq)) ||
        (mode == PQ_SEARCH_LEARNED &&
         (rc = pq_slot(ctx, SLOT_SEEDS, batch * 4 * sizeof(SeedT<R>), &seeds)))) {
        cleanup();
        return rc;
    }
    // the seeded search walks the seed kernel's work list with a grid that fills
    // the GPU once (as many blocks per compute unit as fit at once)
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
            hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_find_list<R>, 256, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    const uint64_t list_grid = (uint64_t) cus * per_cu;
    // the work list of the seeded search (the plain search's list kernel walks every
    // chain in order and only uses the counters)
    if (mode == PQ_SEARCH_LEARNED) {
        cb.wq_cnt = cb.wq + batch * 4;
    } else {
        cb.wq_cnt = cb.wq;
        cb.wq = nullptr;
    }
    ExtOut eo;
    eo.ovB = ctx->d_ovB;
    eo.ovF = ctx->d_ovF;
    eo.numB = ctx->d_numB;
    eo.numF = ctx->d_numF;
    eo.recB = rec;
    eo.recF = recF;
    eo.edges = app;
    eo.staged = staged;
    eo.edge_cap = edge_cap;
    eo.edge_count = (unsigned long long *) misc;          // misc[0..1]
    eo.n_overflow = misc + 2;
    eo.flags = misc + 3;
    eo.overflow_reads = ovl;
    // work counters: spread copies, summed into misc[8..39] at the end
    unsigned long long *fstats = wstats;
    eo.xstats = wstats;
    eo.tier = 0;
    // export during the search (eager only): the batch's scan input and output,
    // node records and the two words of enqueue_node_records' layout, two staging
    // buffers of edge_t records used in turn, and the events ordering the
    // conversions on the search stream against the copies on the transfer stream
    uint64_t *xtot = nullptr, *xscan = nullptr;
    unsigned long long *xwords = nullptr;
    uint2 *xrecs = nullptr;
    uint8_t *xstage[2] = {nullptr, nullptr};
    size_t xstage_bytes[2] = {0, 0};
    hipEvent_t xconv[2] = {nullptr, nullptr}, xdone[2] = {nullptr, nullptr};
    bool xdone_set[2] = {false, false};
    uint64_t x_lo = 0;          // edges of the batches exported so far
    bool x_stop = false;        // an error flag or a full edge buffer: no more edge copies
    const size_t x_rec = xp && xp->index_bytes == 4 ? 8 : 16;
    auto xcleanup = [&]() {
        if (ctx->xstream)
            (void) hipStreamSynchronize(ctx->xstream);
        for (int i = 0; i < 2; i++) {
            if (xconv[i])
                (void) hipEventDestroy(xconv[i]);
            if (xdone[i])
                (void) hipEventDestroy(xdone[i]);
        }
    };
    if (xp) {
        xp->len_sum = 0;
        xp->edges_short = false;
        uint8_t *nb = nullptr;
        if (!ctx->xstream &&
            hipStreamCreateWithFlags(&ctx->xstream, hipStreamNonBlocking) != hipSuccess) {
            ctx->xstream = nullptr;
            cleanup();
            return ctx->fail(PQ_ERR_HIP, "transfer stream");
        }
        if ((rc = pq_slot(ctx, SLOT_XTOT, (batch + 2) * 8, &xtot)) ||
            (rc = pq_slot(ctx, SLOT_XSCAN, (batch + 2) * 8, &xscan)) ||
            (rc = pq_slot(ctx, SLOT_NODES, n * 8 + 16, &nb))) {
            cleanup();
            return rc;
        }
        xwords = (unsigned long long *) nb;
        xrecs = (uint2 *) (nb + 16);
        for (int i = 0; i < 2; i++)
            if (hipEventCreateWithFlags(&xconv[i], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&xdone[i], hipEventDisableTiming) != hipSuccess) {
                xcleanup();
                cleanup();
                return ctx->fail(PQ_ERR_HIP, "transfer events");
            }
    }
    // the whole search as a sequence of stream operations (captured into a graph
    // in graph mode: kernels only, no memsets, no host synchronisation)
    // search kernel: the autotuner's or the caller's choice, else by the size of
    // both indexes' occurrence lines (the autotuner's picks on this GPU: two
    // chains per lane while the lines fit the last-level cache, one chain per
    // lane up to a few GiB, quad-loaded lines beyond, where scattered loads of
    // whole lines by single lanes lose half their rate)
    const uint64_t occ_bytes = (ctx->dir[0].nlines + ctx->dir[1].nlines) * sizeof(OccLine);
    const uint32_t chains = ctx->find_chains >= 1 && ctx->find_chains <= 4 ? ctx->find_chains
                            : occ_bytes < (1ull << 30)                     ? 2u
                            : occ_bytes < (4ull << 30)                     ? 1u
                                                                           : 3u;
    auto enqueue = [&](bool timed) -> int {
    int rc = PQ_OK;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, misc, 48u);
    hipLaunchKernelGGL(k_wstat_clear, dim3(WSTAT_SLOTS), dim3(64), 0, ctx->stream, wstats,
                       (uint32_t) WSTAT_WORDS);
    const uint32_t shape = ctx->ext_shape < EXT_NSHAPES ? ctx->ext_shape : 0;
    const uint32_t kmax1 = EXT_SHAPES[shape][1];
    // extraction tiers: the first (the tuned shape) over every read of the batch,
    // then each later tier over the previous tier's overflow list, read with its
    // count from the device (no host round trip): (items, lanes per read, waves
    // per block); the one-wave tiers hold a read per block, so that the LDS of a
    // 256-item read does not keep a 512-item read's share from other waves
    struct TierShape {
        uint32_t kmax, gw, waves;
    };
    // items = lanes: the register tiers (one item per lane); then LDS tiers
    static const TierShape TIERS[5] = {
        {16, 4, 4}, {32, 8, 4}, {64, 16, 4}, {256, 64, 1}, {512, 64, 1}};
    uint32_t first_tier = 0;
    while (first_tier < 2 && TIERS[first_tier].kmax < kmax1)
        first_tier++;
    const uint32_t ntiers = 5 - first_tier;
    // tier i writes its overflow list at ovl + i * batch and its count at misc[40 + i]
    uint32_t *const tier_cnt = misc + 40;
    ExtOut eoT[5];
    for (uint32_t i = 0; i < ntiers; i++) {
        eoT[i] = eo;
        eoT[i].overflow_reads = ovl + i * batch;
        eoT[i].n_overflow = tier_cnt + i;
        // work counters per tier class: the first, the register ones after it,
        // the LDS ones
        const uint32_t kmx = TIERS[first_tier + i].kmax;
        eoT[i].tier = i == 0 ? 0u : kmx <= 64 ? 1u : 2u;
    }
    ExtOut eoG = eo;
    eoG.tier = 3;
    // key words of the register tiers: 1, 2, 3, 4 or 8 (zero words past a key)
    const uint32_t kwr = kw <= 4 ? kw : 8;
    auto launch_tier = [&](uint32_t i, uint64_t r0, uint64_t nr) {
        const TierShape ts = TIERS[first_tier + i];
        const uint32_t groups = ts.waves * 64 / ts.gw;
        const uint32_t *list = i ? eoT[i - 1].overflow_reads : nullptr;
        const uint32_t *cnt = i ? eoT[i - 1].n_overflow : nullptr;
        const dim3 grid(i ? (uint32_t) std::min<uint64_t>(pq_blocks(batch, groups), 4096)
                          : pq_blocks(nr, groups));
        const dim3 block(64 * ts.waves);
        if (ts.kmax <= 64 && kwr <= 4) {
#define PQ_REG_TIER(GG, KK)                                                                    \
    hipLaunchKernelGGL((k_extract_reg<GG, KK, R>), grid, block, 0, ctx->stream, G, fw, bw, cb, r0, \
                       nr, list, cnt, eoT[i])
#define PQ_REG_TIERS(GG)                                                                       \
    switch (kwr) {                                                                             \
    case 1: PQ_REG_TIER(GG, 1); break;                                                         \
    case 2: PQ_REG_TIER(GG, 2); break;                                                         \
    case 3: PQ_REG_TIER(GG, 3); break;                                                         \
    case 4: PQ_REG_TIER(GG, 4); break;                                                         \
    default: PQ_REG_TIER(GG, 4);                                                               \
    }
            if (ts.kmax == 16) {
                PQ_REG_TIERS(4)
            } else if (ts.kmax == 32) {
                PQ_REG_TIERS(8)
            } else {
                PQ_REG_TIERS(16)
            }
#undef PQ_REG_TIERS
#undef PQ_REG_TIER
            return;
        }
        // keys of more than 4 words: the small tiers sort in LDS
        if (ts.kmax <= 64) {
            const uint32_t gw = ts.kmax / 2, gl = EXT_WAVES * 64 / gw;
            const size_t ldsl = ext_wave_bytes(ts.kmax, kw, true) * gl;
            const dim3 gridl(i ? (uint32_t) std::min<uint64_t>(pq_blocks(batch, gl), 4096)
                               : pq_blocks(nr, gl));
            const dim3 blockl(64 * EXT_WAVES);
            if (ts.kmax == 16)
                hipLaunchKernelGGL((k_extract_lds<16, EXT_WAVES, 8, R>), gridl, blockl, ldsl,
                                   ctx->stream, G, fw, bw, cb, r0, nr, list, cnt, kw, eoT[i]);
            else if (ts.kmax == 32)
                hipLaunchKernelGGL((k_extract_lds<32, EXT_WAVES, 16, R>), gridl, blockl, ldsl,
                                   ctx->stream, G, fw, bw, cb, r0, nr, list, cnt, kw, eoT[i]);
            else
                hipLaunchKernelGGL((k_extract_lds<64, EXT_WAVES, 32, R>), gridl, blockl, ldsl,
                                   ctx->stream, G, fw, bw, cb, r0, nr, list, cnt, kw, eoT[i]);
            return;
        }
        const size_t lds = ext_wave_bytes(ts.kmax, kw) * groups;
        if (ts.kmax == 256)
            hipLaunchKernelGGL((k_extract_lds<256, 1, 64, R>), grid, block, lds, ctx->stream, G, fw,
                               bw, cb, r0, nr, list, cnt, kw, eoT[i]);
        else
            hipLaunchKernelGGL((k_extract_lds<KMAX_LDS2, 1, 64, R>), grid, block, lds, ctx->stream,
                               G, fw, bw, cb, r0, nr, list, cnt, kw, eoT[i]);
    };
    // per batch: seeds, candidates, then the extraction tiers; the later tiers
    // take their read lists and counts from the device (no host round trip)
    for (uint64_t b = 0; b < nbatch && rc == PQ_OK; b++) {
        const uint64_t r0 = bnd[b], nr = bnd[b + 1] - bnd[b];
        hipEvent_t *e = &ev[5 * b];
        uint32_t kseed = 0;
        if (timed)
            (void) hipEventRecord(e[0], ctx->stream);
        const bool listed = mode == PQ_SEARCH_LEARNED || chains == 4;
        if (listed)
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, cb.wq_cnt, 2u);
        if (mode == PQ_SEARCH_LEARNED &&
            (rc = pq_learned_seeds(ctx, r0, nr, seeds, &kseed, cb, fstats,
                                   timed ? e[4] : nullptr)))
            break;
        if (timed)
            (void) hipEventRecord(e[1], ctx->stream);
        if (listed)
            hipLaunchKernelGGL(k_find_list<R>,
                               dim3((uint32_t) std::min<uint64_t>(pq_blocks(nr * 4, 256),
                                                                  list_grid)),
                               dim3(256), 0, ctx->stream, G, fw, bw, r0, (uint32_t) (nr * 4), t,
                               Ct, seeds, kseed, cb, fstats);
        else if (chains == 1)
            hipLaunchKernelGGL(k_find1<R>, dim3(pq_blocks(nr * 4, 256)), dim3(256), 0,
                               ctx->stream, G, fw, bw, r0, nr, t, Ct, seeds, kseed, cb, fstats);
        else if (chains == 3)
            hipLaunchKernelGGL(k_find1q<R>, dim3(pq_blocks(nr * 4, 256)), dim3(256), 0,
                               ctx->stream, G, fw, bw, r0, nr, t, Ct, seeds, kseed, cb, fstats);
        else
            hipLaunchKernelGGL(k_find2<R>, dim3(pq_blocks(nr * 2, 256)), dim3(256), 0,
                               ctx->stream, G, fw, bw, r0, nr, t, Ct, seeds, kseed, cb, fstats);
        if (cb.vmax)
            hipLaunchKernelGGL(k_verify<R>, dim3(pq_blocks(nr * 4, 256)), dim3(256), 0,
                               ctx->stream, G, fw, bw, cb, r0, nr, t, fstats);
        if (timed)
            (void) hipEventRecord(e[2], ctx->stream);
        for (uint32_t i = 0; i < ntiers; i++)
            launch_tier(i, r0, nr);
        // sides beyond the global tier: listed (count misc[7]) and sorted by the whole
        // GPU right here, before the next batch reuses the candidate buffers; a
        // graph capture cannot wait for the count, so there they raise flag 2
        hipLaunchKernelGGL(k_extract_global<R>, dim3(GLOBAL_SLOTS), dim3(64), 0, ctx->stream, G, fw,
                           bw, cb, r0, eoT[ntiers - 1].overflow_reads, eoT[ntiers - 1].n_overflow,
                           kw, gscratch, eoG, huge,
                           misc + 7, timed ? 0u : 1u);
        PQ_CHECK(hipGetLastError());
        if (timed) {
            uint32_t nh = 0;
            PQ_CHECK(hipMemcpyAsync(&nh, misc + 7, 4, hipMemcpyDeviceToHost, ctx->stream));
            PQ_CHECK(hipStreamSynchronize(ctx->stream));
            if (nh) {
                std::vector<HugeRead> hl(nh);
                PQ_CHECK(hipMemcpyAsync(hl.data(), huge, nh * sizeof(HugeRead),
                                        hipMemcpyDeviceToHost, ctx->stream));
                PQ_CHECK(hipStreamSynchronize(ctx->stream));
                if ((rc = extract_huge<R>(ctx, G, fw, bw, cb, r0, kw, eo, hl.data(), nh)))
                    break;
            }
        }
        // overflow counters back to zero for the next batch (misc[40..44], misc[7])
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, tier_cnt, 5u);
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, misc + 7, 1u);
        PQ_CHECK(hipGetLastError());
        if (timed)
            (void) hipEventRecord(e[3], ctx->stream);
        if (xp && timed) {
            // this batch's part of the CSR (its reads' edges are final) and its node
            // records
            if (b == 0) {
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream,
                                   (uint32_t *) ctx->d_edge_off, 2u);
                hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, ctx->stream, xwords, 2u);
            }
            hipLaunchKernelGGL(k_batch_totals, dim3(pq_blocks(nr + 2, 256)), dim3(256), 0,
                               ctx->stream, ctx->d_numB + r0, ctx->d_numF + r0, nr,
                               ctx->d_edge_off + r0, xtot);
            if ((rc = pq_scan_u64(ctx, xtot, xscan, nr + 2, nullptr)))
                break;
            PQ_CHECK(hipMemcpyAsync(ctx->d_edge_off + r0, xscan + 1, (nr + 1) * 8,
                                    hipMemcpyDeviceToDevice, ctx->stream));
            hipLaunchKernelGGL(k_csr, dim3(pq_blocks(nr, 256)), dim3(256), 0, ctx->stream,
                               rec + r0, recF + r0, ctx->d_numB + r0, ctx->d_numF + r0,
                               ctx->d_edge_off + r0, staged + r0 * 2 * EDGE_SLOT, app, nr,
                               misc + 3, ctx->d_edges);
            hipLaunchKernelGGL(k_node_records, dim3(pq_blocks_gs(nr, 256)), dim3(256), 0,
                               ctx->stream, ctx->d_ovB + r0, ctx->d_ovF + r0, ctx->d_numB + r0,
                               ctx->d_numF + r0, nr, xrecs + r0, (uint32_t *) (xwords + 1));
            PQ_CHECK(hipGetLastError());
            uint64_t hi = 0;
            uint32_t fl = 0;
            PQ_CHECK(hipMemcpyAsync(&hi, ctx->d_edge_off + r0 + nr, 8, hipMemcpyDeviceToHost,
                                    ctx->stream));
            PQ_CHECK(hipMemcpyAsync(&fl, misc + 3, 4, hipMemcpyDeviceToHost, ctx->stream));
            PQ_CHECK(hipStreamSynchronize(ctx->stream));
            // an error flag leaves the counts beyond the buffers: the search is run
            // again or fails, and nothing more is converted
            if (fl & 3u)
                x_stop = true;
            if (!x_stop && xp->edges && hi > xp->edge_capacity) {
                xp->edges_short = true;
                x_stop = true;
            }
            const int sb = (int) (b & 1);
            const bool conv = !x_stop && xp->edges && hi > x_lo;
            if (conv) {
                const size_t need = (hi - x_lo) * x_rec;
                if (xstage_bytes[sb] < need) {
                    if ((rc = pq_slot(ctx, SLOT_XSTAGE0 + sb, need, &xstage[sb])))
                        break;
                    xstage_bytes[sb] = need;
                }
                // the staging buffer's previous copy (two batches back) must be done
                if (xdone_set[sb])
                    PQ_CHECK(hipStreamWaitEvent(ctx->stream, xdone[sb], 0));
                hipLaunchKernelGGL(k_ref_edges, dim3(pq_blocks(hi - x_lo, 256)), dim3(256), 0,
                                   ctx->stream, ctx->d_edges + x_lo, hi - x_lo, xp->index_bytes,
                                   xstage[sb]);
                PQ_CHECK(hipGetLastError());
            }
            PQ_CHECK(hipEventRecord(xconv[sb], ctx->stream));
            PQ_CHECK(hipStreamWaitEvent(ctx->xstream, xconv[sb], 0));
            if (conv)
                PQ_CHECK(hipMemcpyAsync(xp->edges + x_lo * x_rec, xstage[sb], (hi - x_lo) * x_rec,
                                        hipMemcpyDeviceToHost, ctx->xstream));
            if (xp->nodes)
                PQ_CHECK(hipMemcpyAsync(xp->nodes + r0, xrecs + r0, nr * 8, hipMemcpyDeviceToHost,
                                        ctx->xstream));
            PQ_CHECK(hipEventRecord(xdone[sb], ctx->xstream));
            xdone_set[sb] = true;
            if (!x_stop)
                x_lo = hi;
        }
    }
    // the work counters' copies summed into misc[8..39]
    hipLaunchKernelGGL(k_wstat_sum, dim3(1), dim3(64), 0, ctx->stream, wstats,
                       (uint32_t) WSTAT_WORDS, (unsigned long long *) (misc + 8));
    return rc;
    };
    if (as_graph) {
        rc = run_search_graph(ctx, mode, n, batch, edge_cap, misc, rec, recF, staged, app, tot,
                              enqueue, {(uint64_t) (uintptr_t) cb.lc, (uint64_t) (uintptr_t) cb.o,
                                        (uint64_t) (uintptr_t) cb.n, (uint64_t) (uintptr_t) seeds,
                                        (uint64_t) (uintptr_t) gscratch, (uint64_t) (uintptr_t) ovl,
                                        (uint64_t) t, (uint64_t) kw, (uint64_t) Ct.c[2],
                                        (uint64_t) Ct.c[3], (uint64_t) Ct.c[4], (uint64_t) Ct.c[5],
                                        (uint64_t) sizeof(R), (uint64_t) chains,
                                        (uint64_t) (uintptr_t) G.off, (uint64_t) (uintptr_t) G.ends,
                                        (uint64_t) (uintptr_t) G.end_rank, (uint64_t) G.L,
                                        (uint64_t) G.Lmax, (uint64_t) G.N, (uint64_t) cb.cap,
                                        (uint64_t) (uintptr_t) cb.vl, (uint64_t) (uintptr_t) cb.vjob,
                                        (uint64_t) cb.vmax});
        xcleanup();
        cleanup();
        return rc;
    }
    // the edge array at the bound the graph path uses, so that the whole search,
    // edge compaction included, runs with one host wait at its end
    const uint64_t ub = 2ull * EDGE_SLOT * n + edge_cap + 1;
    if ((rc = pq_slot(ctx, SLOT_EDGES, ub * 8, &ctx->d_edges)) ||
        (!ctx->tmp(SLOT_SCAN, pq_scan_scratch_bytes(n + 1, 8)) &&
         (rc = ctx->fail(PQ_ERR_NOMEM, "scan scratch")))) {
        xcleanup();
        cleanup();
        return rc;
    }
    if (xp && xp->reserve_only) {
        // staging for about four edges per read of a batch (grown on demand)
        for (int i = 0; i < 2 && rc == PQ_OK; i++)
            rc = pq_slot(ctx, SLOT_XSTAGE0 + i, 4 * batch * x_rec, &xstage[i]);
        xcleanup();
        cleanup();
        return rc;
    }
    hipEvent_t ce[2] = {nullptr, nullptr};
    (void) hipEventCreate(&ce[0]);
    (void) hipEventCreate(&ce[1]);
    rc = enqueue(true);
    if (rc == PQ_OK && !xp) {
        (void) hipEventRecord(ce[0], ctx->stream);
        hipLaunchKernelGGL(k_edge_totals, dim3(pq_blocks(n + 1, 256)), dim3(256), 0, ctx->stream,
                           ctx->d_numB, ctx->d_numF, n, tot);
        rc = pq_scan_u64(ctx, tot, ctx->d_edge_off, n + 1, nullptr);
    }
    if (rc == PQ_OK && !xp) {
        hipLaunchKernelGGL(k_csr, dim3(pq_blocks(n, 256)), dim3(256), 0, ctx->stream, rec, recF,
                           ctx->d_numB, ctx->d_numF, ctx->d_edge_off, staged, app, n, misc + 3,
                           ctx->d_edges);
        (void) hipEventRecord(ce[1], ctx->stream);
        if (hipGetLastError() != hipSuccess)
            rc = ctx->fail(PQ_ERR_HIP, "edge compaction launch");
    }
    if (rc == PQ_OK) {
        uint32_t hm[40];
        uint64_t total = 0;
        if (hipMemcpyAsync(hm, misc, sizeof(hm), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(&total, ctx->d_edge_off + n, 8, hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = ctx->fail(PQ_ERR_HIP, "search result read-back");
        if (rc == PQ_OK) {
            for (uint64_t b = 0; b < nbatch; b++) {
                hipEvent_t *e = &ev[5 * b];
                if (mode == PQ_SEARCH_LEARNED) {
                    ctx->timing.learned_ms += ev_ms(e[0], e[1]);
                    ctx->timing.seed_kernel_ms += ev_ms(e[4], e[1]);
                }
                double fms = ev_ms(e[1], e[2]);
                ctx->timing.find_ms += fms;
                ctx->timing.find_kernel_ms += fms;
                ctx->timing.find_launches++;
                ctx->timing.extract_ms += ev_ms(e[2], e[3]);
            }
            ctx->timing.compact_ms = xp ? 0.0 : ev_ms(ce[0], ce[1]);
            search_counters(ctx, hm);
            if (hm[3] & 2u) {
                rc = ctx->fail(PQ_ERR_STATE, "oversized read side left unprocessed");
            } else if (hm[3] & 1u) {
                ctx->edge_need = (uint64_t) hm[0] | ((uint64_t) hm[1] << 32);
                rc = PQ_RETRY_EDGES;
            } else {
                ctx->n_edges = total;
                ctx->searched = true;
            }
        }
    }
    if (rc == PQ_OK && xp) {
        // the length sum over all edges and the 65535-edge flag; then the last
        // batch's copies
        unsigned long long w[2] = {0, 0};
        if (ctx->n_edges)
            hipLaunchKernelGGL(k_edge_len_sum,
                               dim3(std::min<unsigned>(pq_blocks_gs(ctx->n_edges, 256), 2048u)),
                               dim3(256), 0, ctx->stream, ctx->d_edges, ctx->n_edges, xwords);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(w, xwords, 16, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->xstream) != hipSuccess)
            rc = ctx->fail(PQ_ERR_HIP, "graph export");
        else if (w[1])
            rc = ctx->fail(PQ_ERR_LIMIT, "a node has more than 65535 edges on one side");
        else
            xp->len_sum = w[0];
    }
    for (auto &e : ce)
        if (e)
            (void) hipEventDestroy(e);
    xcleanup();
    cleanup();
    return rc;
}

// One search with its retries: a graph replay that met a read side beyond the
// global tier runs again eagerly; a search whose edges overflowed the shared
// pool runs again with the pool at the demand it counted.
static int find_overlaps_retry(pq_ctx_t *ctx, int mode, ExportSpec *xp)
{
    if (!ctx->dir[0].built || !ctx->dir[1].built)
        return ctx->fail(PQ_ERR_STATE, "index not built");
    const bool as_graph = (mode & PQ_SEARCH_GRAPH) != 0;
    mode &= ~PQ_SEARCH_GRAPH;
    if (mode != PQ_SEARCH_PLAIN && mode != PQ_SEARCH_LEARNED)
        return ctx->fail(PQ_ERR_ARG, "unknown search mode");
    if (mode == PQ_SEARCH_LEARNED && !ctx->learned)
        return ctx->fail(PQ_ERR_STATE, "learned index not trained");
    if (xp && as_graph)
        return ctx->fail(PQ_ERR_ARG, "the export during a search needs an eager search");
    auto attempt = [&](bool graph) {
        return ctx->wide ? find_overlaps_t<uint64_t>(ctx, mode, graph, xp)
                         : find_overlaps_t<uint32_t>(ctx, mode, graph, xp);
    };
    bool graph = as_graph;
    int rc = PQ_OK;
    for (int tries = 0; tries < 4; tries++) {
        rc = attempt(graph);
        if (rc == PQ_RETRY_EAGER)
            graph = false;
        else if (rc == PQ_RETRY_EDGES)
            ctx->edge_cap_auto = ctx->edge_need + ctx->edge_need / 8 + 1024;
        else
            break;
    }
    if (rc > 0)
        rc = ctx->fail(PQ_ERR_STATE, "search did not settle its buffers");
    return rc;
}

extern "C" int pq_find_overlaps(pq_ctx_t *ctx, int mode)
{
    if (!ctx)
        return PQ_ERR_ARG;
    return find_overlaps_retry(ctx, mode, nullptr);
}

extern "C" int pq_find_overlaps_export(pq_ctx_t *ctx, int mode, pq_node_t *nodes, void *edges,
                                       uint64_t edge_capacity, uint32_t index_bytes,
                                       uint64_t *n_edges, uint64_t *overlap_len_sum)
{
    if (!ctx || (index_bytes != 4 && index_bytes != 8))
        return PQ_ERR_ARG;
    ExportSpec xp;
    xp.nodes = nodes;
    xp.edges = (uint8_t *) edges;
    xp.edge_capacity = edges ? edge_capacity : 0;
    xp.index_bytes = index_bytes;
    xp.len_sum = 0;
    xp.edges_short = false;
    xp.reserve_only = false;
    int rc = find_overlaps_retry(ctx, mode, &xp);
    if (rc)
        return rc;
    if (n_edges)
        *n_edges = ctx->n_edges;
    if (overlap_len_sum)
        *overlap_len_sum = xp.len_sum;
    if (xp.edges_short)
        return ctx->fail(PQ_ERR_LIMIT, "edge buffer holds " + std::to_string(edge_capacity) +
                                           " records, the graph has " +
                                           std::to_string(ctx->n_edges));
    return PQ_OK;
}

extern "C" int pq_reserve_search(pq_ctx_t *ctx, uint32_t index_bytes)
{
    if (!ctx || (index_bytes != 4 && index_bytes != 8))
        return PQ_ERR_ARG;
    if (!ctx->dir[0].built || !ctx->dir[1].built)
        return ctx->fail(PQ_ERR_STATE, "index not built");
    ExportSpec xp = {};
    xp.index_bytes = index_bytes;
    xp.reserve_only = true;
    // (the previous search's result is dropped, as by any search)
    return ctx->wide ? find_overlaps_t<uint64_t>(ctx, PQ_SEARCH_PLAIN, false, &xp)
                     : find_overlaps_t<uint32_t>(ctx, PQ_SEARCH_PLAIN, false, &xp);
}

extern "C" int pq_set_search_batch(pq_ctx_t *ctx, uint64_t reads)
{
    if (!ctx)
        return PQ_ERR_ARG;
    ctx->search_batch = reads;
    pq_search_graph_free(ctx);
    return PQ_OK;
}

extern "C" int pq_graph_sizes(pq_ctx_t *ctx, uint64_t *n_reads, uint64_t *n_edges)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (!ctx->searched)
        return ctx->fail(PQ_ERR_STATE, "no search result");
    if (n_reads)
        *n_reads = ctx->n;
    if (n_edges)
        *n_edges = ctx->n_edges;
    return PQ_OK;
}

extern "C" int pq_graph_copy(pq_ctx_t *ctx, int32_t *overlaps_B, int32_t *overlaps_F,
                             uint32_t *num_B, uint32_t *num_F, uint64_t *edge_off, uint64_t *dest,
                             uint16_t *len, uint8_t *type)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (!ctx->searched)
        return ctx->fail(PQ_ERR_STATE, "no search result");
    const uint64_t n = ctx->n, E = ctx->n_edges;
    if (overlaps_B && n)
        PQ_CHECK(hipMemcpy(overlaps_B, ctx->d_ovB, n * 4, hipMemcpyDeviceToHost));
    if (overlaps_F && n)
        PQ_CHECK(hipMemcpy(overlaps_F, ctx->d_ovF, n * 4, hipMemcpyDeviceToHost));
    if (num_B && n)
        PQ_CHECK(hipMemcpy(num_B, ctx->d_numB, n * 4, hipMemcpyDeviceToHost));
    if (num_F && n)
        PQ_CHECK(hipMemcpy(num_F, ctx->d_numF, n * 4, hipMemcpyDeviceToHost));
    if (edge_off)
        PQ_CHECK(hipMemcpy(edge_off, ctx->d_edge_off, (n + 1) * 8, hipMemcpyDeviceToHost));
    if ((dest || len || type) && E) {
        std::vector<uint64_t> pk(E);
        PQ_CHECK(hipMemcpy(pk.data(), ctx->d_edges, E * 8, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < E; i++) {
            if (dest)
                dest[i] = pk[i] >> 32;
            if (len)
                len[i] = (uint16_t) ((pk[i] >> 8) & 0xFFFF);
            if (type)
                type[i] = (uint8_t) (pk[i] & 0xFF);
        }
    }
    return PQ_OK;
}

// Autotuner: time the search once per first-tier extraction shape on the
// uploaded reads and keep the fastest (every shape gives the same result; reads
// that do not fit a shape go to the next tier).
extern "C" int pq_autotune(pq_ctx_t *ctx, int mode, pq_tune_t *out)
{
    if (!ctx)
        return PQ_ERR_ARG;
    double best = 1e300;
    uint32_t pick = 1;
    pq_tune_t t;
    memset(&t, 0, sizeof(t));
    // extraction shapes, with the candidate search at two chains per lane
    ctx->find_chains = 2;
    t.find_ms[1] = 1e300;
    for (uint32_t sh = 0; sh < EXT_NSHAPES; sh++) {
        ctx->ext_shape = sh;
        int rc = pq_find_overlaps(ctx, mode);
        if (rc)
            return rc;
        double ms = ctx->timing.extract_ms;
        t.extract_ms[sh] = ms;
        t.find_ms[1] = std::min(t.find_ms[1], ctx->timing.find_kernel_ms);
        if (ms < best) {
            best = ms;
            pick = sh;
        }
    }
    ctx->ext_shape = pick;
    // the candidate search kernel (one chain per lane: more waves in flight, each
    // with one line load outstanding, better where lengths vary or the index is
    // HBM-resident; two per lane; one per lane with the lines of a read's four
    // lanes loaded together, better where the lines span more than the
    // translation caches cover) together with the verification threshold (0: no
    // verification, 8 or 32 rows; verification reads the suffix array, whose rows
    // are scattered over tens of GiB at the largest sizes): every pair timed, the
    // fastest kept (search kernel + verification time).  The seeded search has
    // one search kernel (the work-list walk): only the threshold is timed, with
    // the seed phase (prediction and seed kernel, which finishes the chains whose
    // intervals are within the threshold) counted in.
    // With the seeded search, the seed kernel's threshold is timed with it (every
    // pair of the two).
    static const uint32_t VROWS[3] = {0, 8, 32};
    double best_find = 1e300;
    uint32_t fk = 1, vk = 2, sk = 2;
    for (uint32_t k = 0; k < 4; k++)
        t.find_ms[k] = 1e300;
    for (uint32_t v = 0; v < 3; v++)
        t.verify_ms[v] = t.seed_ms[v] = 1e300;
    const bool seeded = mode == PQ_SEARCH_LEARNED;
    const uint32_t klast = seeded ? 1 : 4;
    const uint32_t seed0 = ctx->seed_rows;
    for (uint32_t k = 1; k <= klast; k++) {
        for (uint32_t v = 0; v < 3; v++) {
            for (uint32_t sd = seeded ? 0 : 2; sd < 3; sd++) {
                ctx->find_chains = k;
                ctx->verify_rows = VROWS[v];
                ctx->seed_rows = seeded ? VROWS[sd] : seed0;
                int rc = pq_find_overlaps(ctx, mode);
                if (rc)
                    return rc;
                const double ms = ctx->timing.find_ms + (seeded ? ctx->timing.learned_ms : 0.0);
                t.find_ms[k - 1] = std::min(t.find_ms[k - 1], ms);
                t.verify_ms[v] = std::min(t.verify_ms[v], ms);
                if (seeded)
                    t.seed_ms[sd] = std::min(t.seed_ms[sd], ms);
                if (ms < best_find) {
                    best_find = ms;
                    fk = k;
                    vk = v;
                    sk = sd;
                }
            }
        }
    }
    ctx->find_chains = seeded ? 0 : fk;
    ctx->verify_rows = VROWS[vk];
    ctx->seed_rows = seeded ? VROWS[sk] : seed0;
    t.find_chains = seeded ? 0 : fk;
    t.verify_rows = VROWS[vk];
    t.seed_rows = ctx->seed_rows;
    t.shape = pick;
    t.group_lanes = EXT_SHAPES[pick][0];
    t.items = EXT_SHAPES[pick][1];
    t.nshapes = EXT_NSHAPES;
    if (out)
        *out = t;
    // leave the search result of the chosen shape in place
    return pq_find_overlaps(ctx, mode);
}

extern "C" int pq_set_tuning(pq_ctx_t *ctx, uint32_t shape)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (shape >= EXT_NSHAPES)
        return ctx->fail(PQ_ERR_ARG, "no such extraction shape");
    ctx->ext_shape = shape;
    return PQ_OK;
}

extern "C" int pq_set_find_chains(pq_ctx_t *ctx, uint32_t chains)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (chains > 4)
        return ctx->fail(PQ_ERR_ARG, "search kernel must be 0 (by index size), 1, 2, 3 or 4");
    ctx->find_chains = chains;
    return PQ_OK;
}

extern "C" int pq_set_seed_verify(pq_ctx_t *ctx, uint32_t rows)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (rows > VERIFY_MAX)
        return ctx->fail(PQ_ERR_ARG, "the seed kernel finishes at most 32 rows");
    ctx->seed_rows = rows;
    return PQ_OK;
}

extern "C" int pq_set_verify(pq_ctx_t *ctx, uint32_t rows)
{
    if (!ctx)
        return PQ_ERR_ARG;
    if (rows > VERIFY_MAX)
        return ctx->fail(PQ_ERR_ARG, "verification takes at most 32 rows");
    ctx->verify_rows = rows;
    return PQ_OK;
}

extern "C" int pq_set_edge_capacity(pq_ctx_t *ctx, uint64_t edges)
{
    if (!ctx)
        return PQ_ERR_ARG;
    ctx->edge_cap_set = edges;
    ctx->edge_cap_auto = 0;
    return PQ_OK;
}

extern "C" int pq_graph_copy_edges(pq_ctx_t *ctx, void *edges, uint32_t index_bytes)
{
    if (!ctx || !edges || (index_bytes != 4 && index_bytes != 8))
        return PQ_ERR_ARG;
    if (!ctx->searched)
        return ctx->fail(PQ_ERR_STATE, "no search result");
    const uint64_t E = ctx->n_edges;
    if (E == 0)
        return PQ_OK;
    const size_t rec = index_bytes == 4 ? 8 : 16;
    uint8_t *d = (uint8_t *) ctx->tmp(SLOT_EXPORT, E * rec);
    if (!d)
        return ctx->fail(PQ_ERR_NOMEM, "edge export scratch");
    hipLaunchKernelGGL(k_ref_edges, dim3(pq_blocks(E, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_edges, E, index_bytes, d);
    PQ_CHECK(hipGetLastError());
    PQ_CHECK(hipMemcpyAsync(edges, d, E * rec, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    return PQ_OK;
}

// Node records and the length sum into the SLOT_NODES scratch: words[0] = sum of
// the overlap lengths, words[1] = flag of a side with more than 65535 edges,
// then one uint2 record per read.
static int enqueue_node_records(pq_ctx_t *ctx, unsigned long long **words, uint2 **recs)
{
    const uint64_t n = ctx->n, E = ctx->n_edges;
    uint8_t *buf;
    int rc = pq_slot(ctx, SLOT_NODES, n * 8 + 16, &buf);
    if (rc)
        return rc;
    *words = (unsigned long long *) buf;
    *recs = (uint2 *) (buf + 16);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, ctx->stream, *words, 2u);
    if (n)
        hipLaunchKernelGGL(k_node_records, dim3(pq_blocks_gs(n, 256)), dim3(256), 0, ctx->stream,
                           ctx->d_ovB, ctx->d_ovF, ctx->d_numB, ctx->d_numF, n, *recs,
                           (uint32_t *) (*words + 1));
    if (E)
        hipLaunchKernelGGL(k_edge_len_sum, dim3(std::min<unsigned>(pq_blocks_gs(E, 256), 2048u)),
                           dim3(256), 0, ctx->stream, ctx->d_edges, E, *words);
    PQ_CHECK(hipGetLastError());
    return PQ_OK;
}

extern "C" int pq_host_alloc(pq_ctx_t *ctx, uint64_t bytes, void **p)
{
    if (!ctx || !p)
        return PQ_ERR_ARG;
    *p = nullptr;
    hipError_t e = hipHostMalloc(p, bytes ? bytes : 16, hipHostMallocDefault);
    if (e != hipSuccess) {
        *p = nullptr;
        (void) hipGetLastError();
        return ctx->fail(PQ_ERR_NOMEM, std::string("hipHostMalloc of ") + std::to_string(bytes) +
                                           " bytes: " + hipGetErrorString(e));
    }
    return PQ_OK;
}

extern "C" void pq_host_free(void *p)
{
    if (p)
        (void) hipHostFree(p);
}

extern "C" int pq_graph_export(pq_ctx_t *ctx, pq_node_t *nodes, void *edges, uint32_t index_bytes,
                               uint64_t *overlap_len_sum)
{
    static_assert(sizeof(pq_node_t) == 8, "pq_node_t is 8 bytes");
    if (!ctx || (edges && index_bytes != 4 && index_bytes != 8))
        return PQ_ERR_ARG;
    if (!ctx->searched)
        return ctx->fail(PQ_ERR_STATE, "no search result");
    const uint64_t n = ctx->n, E = ctx->n_edges;
    unsigned long long *words = nullptr;
    uint2 *recs = nullptr;
    int rc = enqueue_node_records(ctx, &words, &recs);
    if (rc)
        return rc;
    const size_t rec = index_bytes == 4 ? 8 : 16;
    uint8_t *d = nullptr;
    if (edges && E) {
        if ((rc = pq_slot(ctx, SLOT_EXPORT, E * rec, &d)))
            return rc;
        hipLaunchKernelGGL(k_ref_edges, dim3(pq_blocks(E, 256)), dim3(256), 0, ctx->stream,
                           ctx->d_edges, E, index_bytes, d);
        PQ_CHECK(hipGetLastError());
    }
    unsigned long long w[2];
    PQ_CHECK(hipMemcpyAsync(w, words, 16, hipMemcpyDeviceToHost, ctx->stream));
    if (nodes && n)
        PQ_CHECK(hipMemcpyAsync(nodes, recs, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (d)
        PQ_CHECK(hipMemcpyAsync(edges, d, E * rec, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    if (w[1])
        return ctx->fail(PQ_ERR_LIMIT, "a node has more than 65535 edges on one side");
    if (overlap_len_sum)
        *overlap_len_sum = w[0];
    return PQ_OK;
}

extern "C" int pq_graph_nodes(pq_ctx_t *ctx, pq_node_t *nodes, uint64_t *overlap_len_sum)
{
    static_assert(sizeof(pq_node_t) == 8, "pq_node_t is 8 bytes");
    if (!ctx || (!nodes && ctx->n))
        return PQ_ERR_ARG;
    if (!ctx->searched)
        return ctx->fail(PQ_ERR_STATE, "no search result");
    const uint64_t n = ctx->n;
    unsigned long long *words = nullptr;
    uint2 *recs = nullptr;
    int rc = enqueue_node_records(ctx, &words, &recs);
    if (rc)
        return rc;
    unsigned long long w[2];
    PQ_CHECK(hipMemcpyAsync(w, words, 16, hipMemcpyDeviceToHost, ctx->stream));
    if (n)
        PQ_CHECK(hipMemcpyAsync(nodes, recs, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    PQ_CHECK(hipStreamSynchronize(ctx->stream));
    if (w[1])
        return ctx->fail(PQ_ERR_LIMIT, "a node has more than 65535 edges on one side");
    if (overlap_len_sum)
        *overlap_len_sum = w[0];
    return PQ_OK;
}
