// pq_chain.h -- candidate records of the overlap search's chains and the
// verification that finishes a chain on the text (shared by the search kernels
// and the learned seed kernel).
#pragma once
#include "pq_dev.h"

// Candidate records per (read, chain).
struct ChainBuf {
    uint2 *lc;               // ('$' rank in the searched index, count), or (read, 1 |
                             // ENTRY_READ) for a candidate found by verification
    uint16_t *o;             // pattern length
    uint32_t *n;             // entries per (read, chain)
    uint32_t cap;            // entries per chain
    uint64_t *vl;            // per chain: interval start handed to verification
    uint32_t *vjob;          // per chain: rows | depth << 16, 0 = none
    uint32_t vmax;           // largest interval finished by verification (0: none)
    uint32_t *wq;            // seeded search: the chains left to search (batch chain
                             // indices), appended by the seed kernel; nullptr otherwise
    uint32_t *wq_cnt;        // [0] chains listed, [1] chains taken by the search kernel
};

// An entry whose count carries ENTRY_READ names its one read directly.
#define ENTRY_READ 0x80000000u

__device__ __forceinline__ uint32_t entry_count(uint2 e)
{
    return e.y & ~ENTRY_READ;
}

// Read i of an entry of an index with reads_index ri.
__device__ __forceinline__ uint64_t entry_read(uint2 e, uint32_t i, const uint32_t *ri)
{
    return (e.y & ENTRY_READ) ? (uint64_t) e.x : (uint64_t) ri[e.x + i];
}

#define VERIFY_MAX 32        // largest interval a chain hands to verification

// Candidate records of one lane's chain, staged in LDS before they go out.  A
// lane records its candidates over a long run of dependent random reads; written
// one by one, each 8- or 2-byte store reached the L2 alone and its line was
// evicted (by the read stream) before the next store to it, so every record cost
// a write request of its own.  Staged, a chain's records leave in one burst at
// its end (or STAGE_SLOTS at a time).  stg: the lane's first slot; slot s at
// stg[s * STAGE_STRIDE] (one bank-conflict-free plane per slot across the block's
// 256 lanes).  Records: read m (an ENTRY_READ entry) and pattern length o.
#define STAGE_SLOTS 8
#define STAGE_STRIDE 256

struct CandStage {
    uint64_t *stg;
    uint2 *lc;
    uint16_t *oo;
    uint32_t n, flushed;     // records so far; the first `flushed` are in lc/oo

    __device__ __forceinline__ CandStage(uint64_t *s, uint2 *l, uint16_t *o)
        : stg(s), lc(l), oo(o), n(0), flushed(0) {}

    __device__ __forceinline__ void add(uint32_t m, uint32_t o)
    {
        stg[(n - flushed) * STAGE_STRIDE] = (uint64_t) m | ((uint64_t) o << 32);
        n++;
        if (n - flushed == STAGE_SLOTS) {
            for (uint32_t s = 0; s < STAGE_SLOTS; s++) {
                const uint64_t v = stg[s * STAGE_STRIDE];
                lc[flushed + s] = make_uint2((uint32_t) v, 1u | ENTRY_READ);
                oo[flushed + s] = (uint16_t) (v >> 32);
            }
            flushed = n;
        }
    }

    // Keeps the records with o <= olast, in order, in lc/oo; returns their count.
    __device__ __forceinline__ uint32_t finish(uint32_t olast)
    {
        uint32_t k = 0;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t m, o;
            if (i < flushed) {
                m = lc[i].x;
                o = oo[i];
            } else {
                const uint64_t v = stg[(i - flushed) * STAGE_STRIDE];
                m = (uint32_t) v;
                o = (uint32_t) (v >> 32);
            }
            if (o <= olast) {
                if (i >= flushed || k != i) {
                    lc[k] = make_uint2(m, 1u | ENTRY_READ);
                    oo[k] = (uint16_t) o;
                }
                k++;
            }
        }
        return k;
    }
};

// Verification of one chain (one lane).  The chain of read n (Ln codes, text
// start sn) stands at depth o0 <= t with rows [l, l + size) of its searched
// index, size >= minsize, nothing recorded yet.  Row j is the suffix at position
// g of read m, p codes into m (in the searched index's text).  Extending the
// chain by one code moves every row one position left; a row survives while the
// codes left of g equal the chain's next codes, so its reach is o0 + e with e
// the codes matched, at most min(p, Ln - 1 - o0).  The interval at depth o holds
// the rows that reach o, so the backward search would go on while at least
// minsize rows reach the depth (2 for FB and BF, 1 for FF and BB): up to o_last =
// the largest (FF, BB) or second largest (FB, BF) reach.  A row that reaches the
// start of m (e = p) is a '$' row of the interval at depth o0 + p: m is a
// candidate there when t <= o0 + p <= o_last -- exactly the entries the backward
// search records, written here as one read each.
//
// The codes left of g, read leftwards in the searched index's text, are read
// rightwards in the other direction's text X (every read reversed in place),
// from position start(m) + len(m) - p.  The chain's next codes are read n's own
// codes: FB n[Ln-o0-1], n[Ln-o0-2], ... = the backward text from sn + o0; BF
// n[o0], n[o0+1], ... = the forward text from sn + o0; FF and BB their
// complements (flip = all ones).  Both are compared 32 codes at a time.
// Candidates are written to lc/oo (at most size entries, staged in stg as in
// CandStage); returns their count.  Two rows are processed per iteration so that
// two chains of dependent loads (suffix-array row, then the text left of it) are
// in flight.
template <typename R>
__device__ uint32_t verify_chain(const Geom &G, const R *sa, const uint64_t *X, const uint64_t *Pt,
                                 uint64_t flip, uint64_t sn, uint32_t Ln, uint32_t o0, uint64_t l,
                                 uint32_t size, uint32_t minsize, uint32_t t, uint2 *lc,
                                 uint16_t *oo, uint64_t *stg)
{
    const uint32_t jmax = Ln - 1 - o0;
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
    uint32_t t1 = 0, t2 = 0;   // the two largest reaches + 1
    CandStage cs(stg, lc, oo);
    for (uint32_t j = 0; j < size; j += 2) {
        uint64_t g[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
            g[h] = j + h < size ? (uint64_t) sa[l + j + h] : 0ull;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (j + h >= size)
                continue;
            const uint64_t m = pq_read_of(G, g[h]);
            const uint64_t sm = pq_read_start(G, m);
            const uint32_t p = (uint32_t) (g[h] - sm);
            const uint32_t J = p < jmax ? p : jmax;
            const uint64_t xs = sm + pq_read_len(G, m) - p, x0 = xs >> 5;
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
            if (p <= jmax && e == p && o0 + p >= t)   // tentative: o_last not known yet
                cs.add((uint32_t) m, o0 + p);
        }
    }
    const uint32_t top = minsize == 1 ? t1 : t2;   // >= o0 + 1: the chain was live at o0
    const uint32_t olast = top - 1 < Ln - 1 ? top - 1 : Ln - 1;
    return cs.finish(olast);
}
