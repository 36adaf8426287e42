// pq_ctx.h -- library context (internal).  One context = one device, one HIP
// stream, and every device buffer of the index and of the last search.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/pasqual_amd.h"
#include "pq_dev.h"

#define PQ_CHECK(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return ctx->fail(PQ_ERR_HIP, std::string(#call) + ": " +           \
                                             hipGetErrorString(e_));           \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Per-direction device index.
struct DirIndex {
    uint64_t *text2 = nullptr;     // packed direction text
    void *sa = nullptr;            // uint32_t or uint64_t positions (pq_ctx::wide)
    OccLine *occ = nullptr;
    uint32_t *reads_index = nullptr;
    uint64_t nlines = 0;
    uint64_t cf[5] = {0, 0, 0, 0, 0};
    uint8_t *bwt8 = nullptr;       // BWT symbol per row, between the SA and occ builds
    bool built = false;
    // sampled keys of the learned index in scratch slot SLOT_SKEY0 + direction:
    // key of every keys_stride-th row at depth keys_D (0: none), written by the
    // last build (pq_set_learned_keys) or training
    uint32_t keys_stride = 0, keys_D = 0;
    uint32_t keys_pending = 0;     // stride of keys the running build writes (0: none)
};

struct Learned;                    // pq_learned.hip

struct pq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t xstream = nullptr; // transfers to the host beside the search (created on use)
    std::string err;

    // reads
    uint64_t n = 0, N = 0;
    uint32_t L = 0, Lmax = 0, Lmin = 0, t = 0;
    std::vector<uint64_t> h_off;   // host copy of offsets
    uint64_t C[6] = {0}, sfx[4] = {0}, pfx[4] = {0};
    uint64_t *d_off = nullptr;
    uint64_t *d_ends = nullptr;    // end-of-read bitmap
    uint32_t *d_end_rank = nullptr;
    // rows and positions: 64-bit when the text has 2^32 codes or more (or when
    // asked for); suffixes are sorted in slices of at most slice_cap
    bool wide = false;
    uint32_t opt_row_bytes = 0;    // 0 automatic, 4 or 8
    uint64_t opt_slice = 0;        // 0 automatic (2^31)
    uint32_t last_slices = 0;      // slices of the last suffix-array build
    bool building = false;         // inside pq_build_index
    bool build_scratch_held = false;   // build scratch slots still allocated

    DirIndex dir[2];

    // search result (device)
    int32_t *d_ovB = nullptr, *d_ovF = nullptr;
    uint32_t *d_numB = nullptr, *d_numF = nullptr;
    uint64_t *d_edge_off = nullptr;
    uint64_t *d_edges = nullptr;   // packed (dest<<32 | len<<8 | type)
    uint64_t n_edges = 0;
    bool searched = false;

    Learned *learned = nullptr;
    uint32_t ext_shape = 0;        // first extraction tier: index into the tuned shapes
    uint32_t find_chains = 0;      // candidate-search kernel (1, 2, 3; 0: by index size)
    uint64_t edge_cap_set = 0;     // shared edge buffer capacity, 0 automatic
    uint32_t verify_rows = 32;     // chains of at most this many rows at depth <= t are
                                   // verified on the text (0: backward search to the end)
    uint32_t seed_rows = 32;       // seeded search: chains of at most this many rows at
                                   // the seed depth are finished by the seed kernel
    uint64_t edge_cap_auto = 0;    // capacity an overflowing search asked for (these reads)
    uint64_t edge_need = 0;        // edges the last overflowing search counted
    uint64_t search_batch = 0;     // reads per search batch, 0 automatic
    uint32_t key_stride = 0;       // builds also write the learned index's sampled keys
                                   // of every key_stride-th row (a power of two; 0: no)
    uint64_t learned_gen = 0;      // bumped by every learned-index training
    // search captured as a graph, valid while its key (launch arguments) holds
    hipGraph_t search_graph = nullptr;
    hipGraphExec_t search_exec = nullptr;
    std::vector<uint64_t> search_key;
    pq_timing_t timing{};

    std::vector<DevBuf> scratch;   // reusable temporaries, by slot

    int fail(int code, const std::string &msg)
    {
        err = msg;
        return code;
    }
    Geom geom() const
    {
        Geom g;
        g.n = n;
        g.N = N;
        g.L = L;
        g.Lmax = Lmax;
        g.off = d_off;
        g.ends = d_ends;
        g.end_rank = d_end_rank;
        return g;
    }
    template <typename R> DirViewT<R> view(int d) const
    {
        DirViewT<R> v;
        v.text2 = dir[d].text2;
        v.occ = dir[d].occ;
        v.reads_index = dir[d].reads_index;
        v.sa = (const R *) dir[d].sa;
        for (int c = 0; c < 5; c++)
            v.cf[c] = (R) dir[d].cf[c];
        return v;
    }
    template <typename R> CTab<R> ctab() const
    {
        CTab<R> t;
        for (int c = 0; c < 6; c++)
            t.c[c] = (R) C[c];
        return t;
    }
    // Scratch slot `slot` with at least `bytes` bytes (grown, never shrunk).
    void *tmp(int slot, size_t bytes);
};

// Scratch slots of pq_ctx::tmp (grow-only device buffers kept for the life of
// the context, so repeated builds and searches allocate nothing).
enum PqSlot {
    SLOT_SCAN = 0, SLOT_RADIX, SLOT_OCC, SLOT_RUNS, SLOT_EXPORT, SLOT_BIGK, SLOT_BIGV,
    SLOT_BK = 10, SLOT_BK2, SLOT_V2, SLOT_FLAGS, SLOT_POS, SLOT_STARTS, SLOT_CLS,
    SLOT_COUNTS, SLOT_LISTS, SLOT_SUBLISTS,
    SLOT_CAND_LC = 20, SLOT_CAND_O, SLOT_CAND_N, SLOT_RECB, SLOT_RECF, SLOT_APPEND,
    SLOT_TOT, SLOT_OVL, SLOT_MISC, SLOT_SEEDS, SLOT_GSCRATCH, SLOT_STAGED,
    SLOT_OVB = 34, SLOT_OVF, SLOT_NUMB, SLOT_NUMF, SLOT_EDGE_OFF, SLOT_EDGES,
    SLOT_LEARNED = 40, SLOT_BWT8 = 47, SLOT_SLICE_HIST, SLOT_SLICE_TILES, SLOT_PROBE = 50,
    SLOT_BIGRUNS, SLOT_NODES, SLOT_MISC2, SLOT_HUGE, SLOT_HUGE_LIST, SLOT_VJOB_L, SLOT_VJOB,
    SLOT_WORKQ, SLOT_WSTATS,
    SLOT_XSTAGE0 = 60, SLOT_XSTAGE1, SLOT_XTOT, SLOT_XSCAN, SLOT_SKEY0, SLOT_SKEY1,
};

// Grow-only slot buffer of at least `bytes` bytes into *p, or PQ_ERR_NOMEM.
template <typename T> static inline int pq_slot(pq_ctx_t *ctx, int slot, size_t bytes, T **p)
{
    *p = (T *) ctx->tmp(slot, bytes);
    if (!*p)
        return ctx->fail(PQ_ERR_NOMEM, "device scratch of " + std::to_string(bytes) + " bytes");
    return PQ_OK;
}

// helpers shared across translation units
int pq_alloc(pq_ctx_t *ctx, void **p, size_t bytes);
void pq_free(void *p);
int pq_scan_u32(pq_ctx_t *ctx, const uint32_t *in, uint32_t *out, uint64_t count,
                uint32_t *total_host);
int pq_scan_u64(pq_ctx_t *ctx, const uint64_t *in, uint64_t *out, uint64_t count,
                uint64_t *total_host);
size_t pq_scan_scratch_bytes(uint64_t count, size_t elem);
// one stable 8-bit LSD radix pass of (key, value) pairs by the digit at shift
int pq_radix_pass(pq_ctx_t *ctx, const uint64_t *kin, const uint32_t *vin, uint64_t *kout,
                  uint32_t *vout, uint64_t count, unsigned shift);
int pq_radix_pass(pq_ctx_t *ctx, const uint64_t *kin, const uint64_t *vin, uint64_t *kout,
                  uint64_t *vout, uint64_t count, unsigned shift);
int pq_radix_pass(pq_ctx_t *ctx, const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                  uint32_t *vout, uint64_t count, unsigned shift);
int pq_radix_pass(pq_ctx_t *ctx, const uint32_t *kin, const uint64_t *vin, uint32_t *kout,
                  uint64_t *vout, uint64_t count, unsigned shift);
void pq_search_graph_free(pq_ctx_t *ctx);
int pq_build_sa_dir(pq_ctx_t *ctx, int d);
int pq_build_occ_dir(pq_ctx_t *ctx, int d);
void pq_learned_free(pq_ctx_t *ctx);
// Seed of one (read, chain) of the candidate search at the learned depth:
// search interval [l, l + size); size SEED_NONE makes the chain start from
// depth 1 instead.
template <typename R> struct SeedT {
    R l;
    uint32_t size;
};
#define SEED_NONE 0xFFFFFFFFu
// size SEED_DONE: the seed kernel finished the chain by verification (its
// records are written; the search kernels leave it alone)
#define SEED_DONE 0xFFFFFFFEu
struct ChainBuf;
// d_seeds: SeedT<uint32_t> or SeedT<uint64_t> per (read, chain) as ctx->wide says;
// chains whose interval holds at most cb.vmax rows are finished into cb, the rows
// read counted as WS_SEED_ROWS; `mid` (optional) is recorded between the model's
// prediction and the seed kernel
int pq_learned_seeds(pq_ctx_t *ctx, uint64_t r0, uint64_t nr, void *d_seeds, uint32_t *kseed,
                     const ChainBuf &cb, unsigned long long *stats, hipEvent_t mid);
void pq_release_build_scratch(pq_ctx_t *ctx);

static inline unsigned pq_blocks(uint64_t work, unsigned per)
{
    return (unsigned) ((work + per - 1) / per);
}

// Blocks for a grid-stride kernel over `work` items: at most 2^20 blocks.
static inline unsigned pq_blocks_gs(uint64_t work, unsigned per)
{
    uint64_t b = (work + per - 1) / per;
    return (unsigned) (b < (1u << 20) ? (b ? b : 1) : (1u << 20));
}
