/*
 * pq_oracle.c -- CPU oracle (test infrastructure only; see pq_oracle.h).
 *
 * Restates, in plain single-threaded C, what the reference computes on the hot
 * path: the canonical suffix arrays of the forward and reversed read texts,
 * BWT / occurrence tables / reads_index in the reference host layout, the
 * interval shortcut, and the FM-index overlap search with irreducible-edge
 * extraction.  A second, independent statement of the overlap search on
 * explicit strings (no FM index) cross-checks the first and defines the
 * semantics for reads of unequal length, which the reference rejects
 * (index.c:259-271).
 */
#include "pq_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CAP_EDGES 65535

static uint8_t comp_code(uint8_t c)
{
    return (uint8_t) (5 - c);      /* 1<->4, 2<->3; ntc_table, common.c:13 */
}

static uint64_t read_of(const uint64_t *off, uint64_t n, uint64_t g)
{
    uint64_t lo = 0, hi = n;       /* largest r with off[r] <= g */
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

void orc_dir_text(const uint8_t *text, const uint64_t *off, uint64_t n, int dir,
                  uint8_t *out)
{
    for (uint64_t r = 0; r < n; r++) {
        uint64_t a = off[r], b = off[r + 1];
        for (uint64_t k = a; k < b; k++)
            out[k] = dir ? text[a + (b - 1 - k)] : text[k];
    }
}

/* ---- suffix array ---------------------------------------------------------- */

typedef struct {
    uint64_t key;                  /* 29 chars (2 bits each, padded) + min(rem,30) */
    uint64_t g;
} sa_item_t;

static const uint8_t *g_txt;       /* comparator context (single-threaded) */
static const uint64_t *g_end;

static int sa_cmp(const void *pa, const void *pb)
{
    const sa_item_t *a = pa, *b = pb;
    if (a->key != b->key)
        return a->key < b->key ? -1 : 1;
    if ((a->key & 31) == 30) {     /* both longer than 29: compare the rest */
        uint64_t ra = g_end[a->g] - a->g, rb = g_end[b->g] - b->g;
        uint64_t m = ra < rb ? ra : rb;
        for (uint64_t k = 29; k < m; k++) {
            uint8_t x = g_txt[a->g + k], y = g_txt[b->g + k];
            if (x != y)
                return x < y ? -1 : 1;
        }
        if (ra != rb)
            return ra < rb ? -1 : 1;   /* proper prefix first */
    }
    return a->g < b->g ? -1 : (a->g > b->g);
}

int orc_sa(const uint8_t *dtext, const uint64_t *off, uint64_t n, uint64_t *sa)
{
    uint64_t N = off[n];
    sa_item_t *it = malloc(sizeof(sa_item_t) * (N ? N : 1));
    uint64_t *end = malloc(sizeof(uint64_t) * (N ? N : 1));
    if (!it || !end) {
        free(it);
        free(end);
        return -1;
    }
    for (uint64_t r = 0; r < n; r++)
        for (uint64_t g = off[r]; g < off[r + 1]; g++)
            end[g] = off[r + 1];
    for (uint64_t g = 0; g < N; g++) {
        uint64_t rem = end[g] - g, k = 0;
        for (uint64_t j = 0; j < 29; j++)
            k = (k << 2) | (j < rem ? (uint64_t) (dtext[g + j] - 1) : 0);
        it[g].key = (k << 5) | (rem < 30 ? rem : 30);
        it[g].g = g;
    }
    g_txt = dtext;
    g_end = end;
    qsort(it, N, sizeof(sa_item_t), sa_cmp);
    for (uint64_t i = 0; i < N; i++)
        sa[i] = it[i].g;
    free(it);
    free(end);
    return 0;
}

void orc_bwt(const uint8_t *dtext, const uint64_t *off, uint64_t n,
             const uint64_t *sa, uint8_t *bwt, uint64_t *reads_index)
{
    uint64_t N = off[n], m = 0;
    for (uint64_t i = 0; i < N; i++) {
        uint64_t g = sa[i], r = read_of(off, n, g);
        if (g == off[r]) {
            bwt[i] = 0;
            if (reads_index)
                reads_index[m++] = r;
        } else {
            bwt[i] = dtext[g - 1];
        }
    }
}

void orc_counts(const uint8_t *text, const uint64_t *off, uint64_t n,
                uint64_t C[6], uint64_t suffix_count[4],
                uint64_t prefix_count[4])
{
    uint64_t cnt[4] = { 0, 0, 0, 0 };
    memset(suffix_count, 0, 4 * sizeof(uint64_t));
    memset(prefix_count, 0, 4 * sizeof(uint64_t));
    for (uint64_t r = 0; r < n; r++) {
        if (off[r + 1] == off[r])
            continue;
        prefix_count[text[off[r]] - 1]++;
        suffix_count[text[off[r + 1] - 1] - 1]++;
        for (uint64_t g = off[r]; g < off[r + 1]; g++)
            cnt[text[g] - 1]++;
    }
    C[0] = 0;
    C[1] = 0;
    for (int c = 1; c <= 4; c++)
        C[c + 1] = C[c] + cnt[c - 1];
}

void orc_occ_ref(const uint8_t *bwt, uint64_t N, const uint64_t fix4[4],
                 uint64_t *bitmap, uint64_t *sample)
{
    uint64_t rows = (N + 64) / 64, run[5];
    run[0] = 0;
    for (int c = 0; c < 4; c++)
        run[c + 1] = fix4[c];
    memset(bitmap, 0, 5 * rows * sizeof(uint64_t));
    for (uint64_t w = 0; w < rows; w++) {
        for (int c = 0; c < 5; c++)
            sample[5 * w + c] = run[c];
        for (uint64_t j = 0; j < 64 && 64 * w + j < N; j++) {
            uint8_t c = bwt[64 * w + j];
            run[c]++;
            if (j < 63)
                bitmap[c * rows + w] |= 2ULL << j;
        }
    }
}

/* ---- interval shortcut ----------------------------------------------------- */

int orc_shortcut(const uint8_t *dtext, uint64_t n, uint32_t L, uint32_t t,
                 const uint64_t *sa, const uint64_t C[6], uint64_t *istart,
                 uint64_t *iend, uint64_t *idepth)
{
    uint64_t N = n * (uint64_t) L;
    for (uint64_t r = 0; r < n; r++) {          /* index.c:495-502, 548-555 */
        uint8_t ch = dtext[r * L + L - 1];
        istart[r] = C[ch];
        iend[r] = C[ch + 1] - C[ch];
        idepth[r] = 1;
    }
    if (t < 9 || N == 0)
        return 0;
    uint32_t Pmax = t < L ? t : L;
    uint32_t nP = Pmax >= 9 ? (Pmax - 9) / 3 + 1 : 0;
    uint64_t *isa = malloc(N * sizeof(uint64_t));
    uint16_t *lcp = malloc(N * sizeof(uint16_t));
    uint64_t *rs = malloc(N * sizeof(uint64_t));
    uint64_t *re = malloc(N * sizeof(uint64_t));
    uint8_t *ps1 = calloc(n * (nP ? nP : 1), 1);   /* per (read, P) state */
    if (!isa || !lcp || !rs || !re || !ps1) {
        free(isa);
        free(lcp);
        free(rs);
        free(re);
        free(ps1);
        return -1;
    }
    for (uint64_t i = 0; i < N; i++)
        isa[sa[i]] = i;
    lcp[0] = 0;
    for (uint64_t i = 1; i < N; i++) {          /* LCP, capped at t */
        uint64_t a = sa[i - 1], b = sa[i];
        uint64_t ra = L - a % L, rb = L - b % L, m = ra < rb ? ra : rb, k = 0;
        if (m > t)
            m = t;
        while (k < m && dtext[a + k] == dtext[b + k])
            k++;
        lcp[i] = (uint16_t) k;
    }
    /* Visit rule: a level-9 bucket of size > 1 is sorted by process_slot_1
     * (sufsort.c:594-611).  Inside process_slot_1 at pos (< L), a child bucket
     * of size f > 1 at pos+3 goes to insert_sort if f*(L-pos) < 3500, else to
     * process_slot_1 (sufsort.c:354-376).  Both record every suffix of length
     * exactly pos (<= t) with the bucket's SA range (sufsort.c:177-209,
     * 322-334); the deepest record wins. */
    for (uint32_t Q = 9; Q <= Pmax; Q += 3) {
        uint64_t start = 0;
        for (uint64_t i = 0; i < N; i++) {
            if (i > 0 && lcp[i] < Q)
                start = i;
            rs[i] = start;
        }
        uint64_t endx = N;
        for (uint64_t i = N; i-- > 0;) {
            re[i] = endx;
            if (lcp[i] < Q)
                endx = i;
        }
        for (uint64_t r = 0; r < n; r++) {
            for (uint32_t p = 0; p < nP; p++) {
                uint32_t P = 9 + 3 * p;
                if (P < Q)
                    continue;
                uint64_t i = isa[r * L + (L - P)];
                uint64_t size = re[i] - rs[i];
                uint8_t *st = &ps1[r * nP + p];
                int vis;
                if (Q == 9) {
                    vis = size > 1;
                    *st = (uint8_t) vis;
                } else {
                    vis = *st && (Q - 3) < L && size > 1;
                    *st = (uint8_t) (vis && size * (uint64_t) (L - (Q - 3)) >= 3500);
                }
                if (Q == P && vis && idepth[r] < P) {
                    istart[r] = rs[i];
                    iend[r] = size;
                    idepth[r] = P;
                }
            }
        }
    }
    free(isa);
    free(lcp);
    free(rs);
    free(re);
    free(ps1);
    return 0;
}

/* ---- graph containers -------------------------------------------------------- */

typedef struct {
    uint64_t *dest;
    uint32_t *len;
    uint8_t *type;
    uint64_t size, cap;
} edgevec_t;

static int ev_push(edgevec_t *v, uint64_t d, uint32_t l, uint8_t ty)
{
    if (v->size == v->cap) {
        uint64_t nc = v->cap ? 2 * v->cap : 1024;
        uint64_t *nd = realloc(v->dest, nc * sizeof(uint64_t));
        if (nd)
            v->dest = nd;
        uint32_t *nl = realloc(v->len, nc * sizeof(uint32_t));
        if (nl)
            v->len = nl;
        uint8_t *nt = realloc(v->type, nc);
        if (nt)
            v->type = nt;
        if (!nd || !nl || !nt)
            return -1;
        v->cap = nc;
    }
    v->dest[v->size] = d;
    v->len[v->size] = l;
    v->type[v->size] = ty;
    v->size++;
    return 0;
}

static int graph_alloc(orc_graph_t *g, uint64_t n)
{
    memset(g, 0, sizeof(*g));
    g->n = n;
    g->overlaps_B = calloc(n + 1, sizeof(int64_t));
    g->overlaps_F = calloc(n + 1, sizeof(int64_t));
    g->num_B = calloc(n + 1, sizeof(uint64_t));
    g->num_F = calloc(n + 1, sizeof(uint64_t));
    g->edge_off = calloc(n + 1, sizeof(uint64_t));
    return (g->overlaps_B && g->overlaps_F && g->num_B && g->num_F && g->edge_off) ? 0 : -1;
}

void orc_graph_free(orc_graph_t *g)
{
    free(g->overlaps_B);
    free(g->overlaps_F);
    free(g->num_B);
    free(g->num_F);
    free(g->edge_off);
    free(g->dest);
    free(g->len);
    free(g->type);
    memset(g, 0, sizeof(*g));
}

/* ---- FM-index overlap search (reference algorithm) --------------------------- */

typedef struct {
    uint64_t ll, uu, remain;
    uint32_t len;
    uint8_t type;
} fm_iv_t;

typedef struct {
    const orc_fmindex_t *ix;
    uint64_t rows;
    uint64_t evals;
    fm_iv_t *arena;
    uint64_t cap;
    edgevec_t side;                /* edges of the side being extracted */
    uint64_t self;                 /* read being processed */
    int overflow;
} fm_t;

enum { T_BF = 0, T_FB = 1, T_FF = 2, T_BB = 3 };

/* OCCx, string_graph.c:11-23: fix + count of c before pos, from the sample of
 * pos's 64-block and the bitmap (bit j+1 = position j) shifted left. */
static uint64_t fm_rank(fm_t *f, int d, int c, uint64_t pos)
{
    uint64_t w = pos / 64;
    unsigned off = (unsigned) (pos % 64);
    uint64_t bits = f->ix->bitmap[d][c * f->rows + w] << (63 - off);
    return f->ix->sample[d][5 * w + c] + (uint64_t) __builtin_popcountll(bits);
}

/* OCCx_LT, string_graph.c:26-45: same over all codes below c (incl. '$'). */
static uint64_t fm_rank_lt(fm_t *f, int d, int c, uint64_t pos)
{
    uint64_t w = pos / 64, s = 0, bits = 0;
    unsigned off = (unsigned) (pos % 64);
    for (int k = 0; k < c; k++) {
        s += f->ix->sample[d][5 * w + k];
        bits |= f->ix->bitmap[d][k * f->rows + w];
    }
    return s + (uint64_t) __builtin_popcountll(bits << (63 - off));
}

/* search_1, string_graph.c:54-82: backward step of [l,u] by c. */
static void fm_step(fm_t *f, int d, int c, uint64_t *l, uint64_t *u)
{
    uint64_t vl = *l, vu = *u + 1;
    const uint64_t *C = f->ix->C;
    f->evals += 2;
    *l = C[c] + fm_rank(f, d, c, vl);
    *u = C[c] + fm_rank(f, d, c, vu) - 1;
}

/* search_2, string_graph.c:85-131: bidirectional step; the partner interval
 * start moves by the number of codes below c in [l,u]. */
static void fm_step2(fm_t *f, int d, int c, uint64_t *l, uint64_t *u,
                     uint64_t *ll, uint64_t *uu)
{
    uint64_t vl = *l, vu = *u + 1;
    const uint64_t *C = f->ix->C;
    f->evals += 2;
    *ll = *ll + (fm_rank_lt(f, d, c, vu) - fm_rank_lt(f, d, c, vl));
    uint64_t ol = fm_rank(f, d, c, vl), ou = fm_rank(f, d, c, vu);
    *l = C[c] + ol;
    *u = C[c] + ou - 1;
    *uu = *ll + ou - ol - 1;
}

/* search_3, string_graph.c:134-159: '$' count in [l,u] minus one, unsigned. */
static uint64_t fm_dollars_m1(fm_t *f, int d, uint64_t l, uint64_t u)
{
    f->evals += 2;
    return fm_rank(f, d, 0, u + 1) - fm_rank(f, d, 0, l) - 1;
}

static int fm_reserve(fm_t *f, uint64_t need)
{
    if (need <= f->cap)
        return 0;
    uint64_t nc = f->cap ? f->cap : 4096;
    while (nc < need)
        nc *= 2;
    fm_iv_t *na = realloc(f->arena, nc * sizeof(fm_iv_t));
    if (!na)
        return -1;
    f->arena = na;
    f->cap = nc;
    return 0;
}

/* extract_fw_edges / extract_bw_edges, string_graph.c:386-731.  The interval
 * list of this level lives at arena[base .. base+size); children go after it.
 * side 0 = B (BF, BB), side 1 = F (FB, FF). */
static int fm_extract(fm_t *f, int side, uint64_t base, uint64_t size,
                      uint32_t level, uint32_t maxlen)
{
    const orc_fmindex_t *ix = f->ix;
    const uint64_t *C = ix->C;
    if (maxlen + level == ix->L) {
        int flag = 0;
        for (uint64_t i = 0; i < size; i++) {
            fm_iv_t *e = &f->arena[base + i];
            if (e->len + level != ix->L)
                continue;
            /* F: FF -> fw index/reads_index, FB -> bw; B: BB -> bw, BF -> fw */
            int d = side ? (e->type == T_FF ? 0 : 1) : (e->type == T_BB ? 1 : 0);
            f->evals += 2;
            uint64_t lo = C[0] + fm_rank(f, d, 0, e->ll);
            uint64_t hi = C[0] + fm_rank(f, d, 0, e->uu + 1) - 1;
            if (lo <= hi) {
                for (uint64_t k = lo; k < hi + 1; k++) {
                    uint64_t m = ix->reads_index[d][k];
                    flag = 1;
                    if (m != f->self) {
                        if (f->side.size + 1 > CAP_EDGES) {
                            f->overflow = 1;
                            return -1;
                        }
                        if (ev_push(&f->side, m, e->len, e->type))
                            return -1;
                    }
                }
                break;
            }
        }
        if (flag)
            return 0;
    }
    uint64_t cbase = base + size;
    for (int a = 1; a <= 4; a++) {
        int b = comp_code((uint8_t) a);
        uint64_t nchild = 0;
        uint32_t mlen = 0;
        if (fm_reserve(f, cbase + size + 1))
            return -1;
        for (uint64_t i = 0; i < size; i++) {
            fm_iv_t *e = &f->arena[base + i];
            if (e->remain == 0)
                continue;
            uint64_t ll = e->ll, uu = e->uu;
            if (side) {
                if (e->type == T_FF)
                    fm_step(f, 0, b, &ll, &uu);
                else
                    fm_step(f, 1, a, &ll, &uu);
            } else {
                if (e->type == T_BB)
                    fm_step(f, 1, a, &ll, &uu);
                else
                    fm_step(f, 0, b, &ll, &uu);
            }
            e->remain -= (uu - ll + 1);
            if (ll <= uu) {
                fm_iv_t *ch = &f->arena[cbase + nchild++];
                ch->ll = ll;
                ch->uu = uu;
                ch->len = e->len;
                ch->type = e->type;
                ch->remain = uu - ll + 1;
                if (mlen < e->len)
                    mlen = e->len;
            }
        }
        if (nchild == 0)
            continue;
        if (fm_extract(f, side, cbase, nchild, level + 1, mlen))
            return -1;
    }
    return 0;
}

int orc_search_fm(const orc_fmindex_t *ix, orc_graph_t *out)
{
    fm_t f;
    memset(&f, 0, sizeof(f));
    f.ix = ix;
    f.rows = (ix->N + 64) / 64;
    if (graph_alloc(out, ix->n))
        return -1;
    edgevec_t all = { 0 };
    const uint64_t *C = ix->C;
    const int L = (int) ix->L, t = (int) ix->t;
    fm_iv_t *Fl = NULL, *Bl = NULL;
    uint64_t capFB = 0;
    int rc = 0;

    for (uint64_t n = 0; n < ix->n && rc == 0; n++) {
        const uint8_t *rd = ix->reads + n * (uint64_t) L;
        uint64_t nF = 0, nB = 0, l, u, ll = 0, uu = 0, tmp;
        int64_t olpF = 0, olpB = 0;
        int i, k, ch;
        f.self = n;
        if (capFB < (uint64_t) (4 * L + 8)) {
            capFB = 4 * (uint64_t) L + 8;
            Fl = realloc(Fl, capFB * sizeof(fm_iv_t));
            Bl = realloc(Bl, capFB * sizeof(fm_iv_t));
            if (!Fl || !Bl) {
                rc = -1;
                break;
            }
        }
        uint64_t ev0 = f.evals;
#define PUSH(list, cnt, ll_, tmp_, len_, ty_)                                  \
        do {                                                                   \
            fm_iv_t *e_ = &(list)[(cnt)++];                                    \
            e_->ll = (ll_);                                                    \
            e_->uu = (ll_) + (tmp_);                                           \
            e_->len = (uint32_t) (len_);                                       \
            e_->type = (ty_);                                                  \
            e_->remain = (tmp_) + 1;                                           \
        } while (0)
        /* F-B: string_graph.c:202-243 */
        l = ix->istart[0][n];
        u = l + ix->iend[0][n] - 1;
        i = L - (int) ix->idepth[0][n] - 1;
        while (l < u && i > L - t - 1) {
            fm_step(&f, 0, rd[i], &l, &u);
            i--;
        }
        if (l < u) {
            ch = rd[i + 1];
            ll = C[ch];
            uu = C[ch + 1] - 1;
            for (k = i + 2; k < L; k++)
                fm_step(&f, 1, rd[k], &ll, &uu);
        }
        while (l < u && i >= 0) {
            tmp = fm_dollars_m1(&f, 0, l, u);
            PUSH(Fl, nF, ll, tmp, L - i - 1, T_FB);
            olpF += (int64_t) (tmp + 1);
            fm_step2(&f, 0, rd[i], &l, &u, &ll, &uu);
            i--;
        }
        /* F-F: string_graph.c:245-286 */
        ch = comp_code(rd[L - 1]);
        l = C[ch];
        u = C[ch + 1] - 1;
        i = L - 2;
        while (l <= u && i > L - t - 1) {
            fm_step(&f, 1, comp_code(rd[i]), &l, &u);
            i--;
        }
        if (l <= u) {
            ch = comp_code(rd[i + 1]);
            ll = C[ch];
            uu = C[ch + 1] - 1;
            for (k = i + 2; k < L; k++)
                fm_step(&f, 0, comp_code(rd[k]), &ll, &uu);
        }
        while (l <= u && i >= 0) {
            tmp = fm_dollars_m1(&f, 1, l, u);
            if (nF >= capFB) {
                rc = -1;
                break;
            }
            PUSH(Fl, nF, ll, tmp, L - i - 1, T_FF);
            olpF += (int64_t) (tmp + 1);
            fm_step2(&f, 1, comp_code(rd[i]), &l, &u, &ll, &uu);
            i--;
        }
        /* B-F: string_graph.c:288-328 */
        l = ix->istart[1][n];
        u = l + ix->iend[1][n] - 1;
        i = (int) ix->idepth[1][n];
        while (l < u && i < t) {
            fm_step(&f, 1, rd[i], &l, &u);
            i++;
        }
        if (l < u) {
            ch = rd[i - 1];
            ll = C[ch];
            uu = C[ch + 1] - 1;
            for (k = i - 2; k >= 0; k--)
                fm_step(&f, 0, rd[k], &ll, &uu);
        }
        while (l < u && i < L) {
            tmp = fm_dollars_m1(&f, 1, l, u);
            PUSH(Bl, nB, ll, tmp, i, T_BF);
            olpB += (int64_t) (tmp + 1);
            fm_step2(&f, 1, rd[i], &l, &u, &ll, &uu);
            i++;
        }
        /* B-B: string_graph.c:330-374 */
        ch = comp_code(rd[0]);
        l = C[ch];
        u = C[ch + 1] - 1;
        i = 1;
        while (l <= u && i < t) {
            fm_step(&f, 0, comp_code(rd[i]), &l, &u);
            i++;
        }
        if (l <= u) {
            ch = comp_code(rd[i - 1]);
            ll = C[ch];
            uu = C[ch + 1] - 1;
            for (k = i - 2; k >= 0; k--)
                fm_step(&f, 1, comp_code(rd[k]), &ll, &uu);
        }
        while (l <= u && i < L) {
            if (i >= t) {
                tmp = fm_dollars_m1(&f, 0, l, u);
                if (nB >= capFB) {
                    rc = -1;
                    break;
                }
                PUSH(Bl, nB, ll, tmp, i, T_BB);
                olpB += (int64_t) (tmp + 1);
            }
            fm_step2(&f, 0, comp_code(rd[i]), &l, &u, &ll, &uu);
            i++;
        }
#undef PUSH
        if (rc)
            break;
        out->rank_evals_find += f.evals - ev0;
        out->overlaps_B[n] = (int32_t) olpB;     /* int in the reference */
        out->overlaps_F[n] = (int32_t) olpF;
        out->edge_off[n] = all.size;
        /* B side first (string_graph.c:823-825), then F (827-829) */
        for (int side = 0; side < 2 && rc == 0; side++) {
            int64_t olp = side ? olpF : olpB;
            uint64_t cnt = side ? nF : nB;
            f.side.size = 0;
            if ((int32_t) olp != 0) {
                if (fm_reserve(&f, cnt + 1)) {
                    rc = -1;
                    break;
                }
                memcpy(f.arena, side ? Fl : Bl, cnt * sizeof(fm_iv_t));
                if (fm_extract(&f, side, 0, cnt, 0, 0))
                    rc = -1;
            }
            for (uint64_t e = 0; e < f.side.size && rc == 0; e++)
                if (ev_push(&all, f.side.dest[e], f.side.len[e], f.side.type[e]))
                    rc = -1;
            if (side)
                out->num_F[n] = f.side.size;
            else
                out->num_B[n] = f.side.size;
        }
    }
    out->edge_off[ix->n] = all.size;
    out->dest = all.dest;
    out->len = all.len;
    out->type = all.type;
    out->rank_evals = f.evals;
    free(f.arena);
    free(f.side.dest);
    free(f.side.len);
    free(f.side.type);
    free(Fl);
    free(Bl);
    return rc;
}

/* ---- explicit-string restatement ------------------------------------------ */

typedef struct {
    uint64_t key;                  /* first K codes of the string, 2 bits each */
    uint64_t m;                    /* read id */
    uint8_t rc;                    /* 0: the read itself, 1: its reverse complement */
} seed_t;

static int seed_cmp(const void *pa, const void *pb)
{
    const seed_t *a = pa, *b = pb;
    if (a->key != b->key)
        return a->key < b->key ? -1 : 1;
    if (a->m != b->m)
        return a->m < b->m ? -1 : 1;
    return (int) a->rc - (int) b->rc;
}

typedef struct {
    const uint8_t *key;            /* extension string */
    uint32_t klen;
    uint32_t o;
    uint64_t m;
    uint8_t rank;                  /* candidate-order rank of the type on its side */
    uint8_t type;
} item_t;

static int item_cmp(const void *pa, const void *pb)
{
    const item_t *a = pa, *b = pb;
    uint32_t mn = a->klen < b->klen ? a->klen : b->klen;
    int c = memcmp(a->key, b->key, mn);
    if (c)
        return c;
    if (a->klen != b->klen)
        return a->klen < b->klen ? -1 : 1;
    if (a->rank != b->rank)
        return (int) a->rank - (int) b->rank;
    if (a->o != b->o)
        return a->o < b->o ? -1 : 1;
    return a->m < b->m ? -1 : (a->m > b->m);
}

static uint64_t pack_k(const uint8_t *s, uint32_t K)
{
    uint64_t k = 0;
    for (uint32_t j = 0; j < K; j++)
        k = (k << 2) | (uint64_t) (s[j] - 1);
    return k;
}

int orc_search_explicit(const uint8_t *text, const uint64_t *off, uint64_t n,
                        uint32_t t, orc_graph_t *out)
{
    uint64_t N = off[n];
    uint32_t K = t < 32 ? t : 32;
    if (graph_alloc(out, n))
        return -1;
    uint8_t *rct = malloc(N ? N : 1);
    seed_t *seeds = malloc(sizeof(seed_t) * (2 * n + 1));
    uint64_t ns = 0, icap = 1024;
    item_t *items = malloc(sizeof(item_t) * icap);
    edgevec_t all = { 0 };
    if (!rct || !seeds || !items)
        goto fail;
    for (uint64_t r = 0; r < n; r++)
        for (uint64_t g = off[r]; g < off[r + 1]; g++)
            rct[g] = comp_code(text[off[r] + (off[r + 1] - 1 - g)]);
    for (uint64_t m = 0; m < n; m++) {
        if (off[m + 1] - off[m] <= t)
            continue;              /* needs o < L_m with o >= t */
        seeds[ns].key = pack_k(text + off[m], K);
        seeds[ns].m = m;
        seeds[ns++].rc = 0;
        seeds[ns].key = pack_k(rct + off[m], K);
        seeds[ns].m = m;
        seeds[ns++].rc = 1;
    }
    qsort(seeds, ns, sizeof(seed_t), seed_cmp);

    for (uint64_t nn = 0; nn < n; nn++) {
        uint32_t Lq = (uint32_t) (off[nn + 1] - off[nn]);
        out->edge_off[nn] = all.size;
        for (int side = 0; side < 2; side++) {   /* 0 = B (query rc(n)), 1 = F (query n) */
            const uint8_t *q = side ? text + off[nn] : rct + off[nn];
            uint64_t ni = 0;
            for (uint32_t o = t; o + 1 <= Lq; o++) {
                if (o == Lq)
                    break;
                const uint8_t *pat = q + (Lq - o);
                uint64_t key = pack_k(pat, K), lo = 0, hi = ns;
                while (lo < hi) {  /* first seed with key >= key */
                    uint64_t mid = (lo + hi) / 2;
                    if (seeds[mid].key < key)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                for (uint64_t s = lo; s < ns && seeds[s].key == key; s++) {
                    uint64_t m = seeds[s].m;
                    uint32_t Lm = (uint32_t) (off[m + 1] - off[m]);
                    const uint8_t *str = seeds[s].rc ? rct + off[m] : text + off[m];
                    if (!(o < Lm) || memcmp(str, pat, o) != 0)
                        continue;
                    if (ni == icap) {
                        icap *= 2;
                        item_t *nit = realloc(items, sizeof(item_t) * icap);
                        if (!nit)
                            goto fail;
                        items = nit;
                    }
                    item_t *it = &items[ni++];
                    it->key = str + o;
                    it->klen = Lm - o;
                    it->o = o;
                    it->m = m;
                    /* F: FB (m) before FF (rc m); B: BF (rc m) before BB (m) */
                    if (side) {
                        it->type = seeds[s].rc ? T_FF : T_FB;
                        it->rank = seeds[s].rc;
                    } else {
                        it->type = seeds[s].rc ? T_BF : T_BB;
                        it->rank = (uint8_t) !seeds[s].rc;
                    }
                }
            }
            if (side)
                out->overlaps_F[nn] = (int64_t) ni;
            else
                out->overlaps_B[nn] = (int64_t) ni;
            qsort(items, ni, sizeof(item_t), item_cmp);
            /* a key is emitted iff no other key is a proper prefix of it; in
             * sorted order that is: the last emitted key is not its prefix */
            const item_t *last = NULL;
            uint64_t emitted = 0;
            for (uint64_t a = 0; a < ni;) {
                uint64_t b = a + 1;
                while (b < ni && items[b].klen == items[a].klen &&
                       memcmp(items[b].key, items[a].key, items[a].klen) == 0)
                    b++;
                int dominated = last && last->klen <= items[a].klen &&
                    memcmp(last->key, items[a].key, last->klen) == 0;
                if (!dominated) {
                    for (uint64_t c = a; c < b; c++) {
                        if (items[c].rank != items[a].rank || items[c].o != items[a].o)
                            break;
                        if (items[c].m == nn)
                            continue;
                        if (ev_push(&all, items[c].m, items[c].o, items[c].type))
                            goto fail;
                        emitted++;
                    }
                    last = &items[a];
                }
                a = b;
            }
            if (emitted > CAP_EDGES)
                goto fail;
            if (side)
                out->num_F[nn] = emitted;
            else
                out->num_B[nn] = emitted;
        }
    }
    out->edge_off[n] = all.size;
    out->dest = all.dest;
    out->len = all.len;
    out->type = all.type;
    free(rct);
    free(seeds);
    free(items);
    return 0;
fail:
    free(rct);
    free(seeds);
    free(items);
    free(all.dest);
    free(all.len);
    free(all.type);
    return -1;
}

/* ---- property check ------------------------------------------------------------ */

int64_t orc_check_sa(const uint8_t *dtext, const uint64_t *off, uint64_t n,
                     const uint64_t *sa)
{
    uint64_t N = off[n];
    uint8_t *seen = calloc(N / 8 + 1, 1);
    uint64_t *end = malloc(sizeof(uint64_t) * (N ? N : 1));
    if (!seen || !end) {
        free(seen);
        free(end);
        return (int64_t) N;
    }
    for (uint64_t r = 0; r < n; r++)
        for (uint64_t g = off[r]; g < off[r + 1]; g++)
            end[g] = off[r + 1];
    int64_t bad = -1;
    for (uint64_t i = 0; i < N && bad < 0; i++) {
        uint64_t g = sa[i];
        if (g >= N || (seen[g / 8] >> (g % 8)) & 1) {
            bad = (int64_t) N;
            break;
        }
        seen[g / 8] |= (uint8_t) (1u << (g % 8));
        if (i == 0)
            continue;
        uint64_t a = sa[i - 1], ra = end[a] - a, rb = end[g] - g, m = ra < rb ? ra : rb, k = 0;
        while (k < m && dtext[a + k] == dtext[g + k])
            k++;
        int ok;
        if (k < m)
            ok = dtext[a + k] < dtext[g + k];
        else if (ra != rb)
            ok = ra < rb;
        else
            ok = a < g;
        if (!ok)
            bad = (int64_t) i;
    }
    free(seen);
    free(end);
    return bad;
}

/* ---- learned index model ------------------------------------------------------ */

double orc_mlp_loss_grad(uint32_t k, uint32_t h, const double *params, const uint64_t *xs,
                         const double *targets, uint64_t n, double *grad)
{
    const double *W1 = params, *b1 = params + (size_t) h * k, *w2 = b1 + h, *b2 = w2 + h;
    const size_t P = (size_t) h * k + 2 * (size_t) h + 1;
    double *f = malloc(sizeof(double) * k), *pre = malloc(sizeof(double) * h);
    double loss = 0;
    if (grad)
        memset(grad, 0, sizeof(double) * P);
    for (uint64_t s = 0; s < n; s++) {
        for (uint32_t j = 0; j < k; j++) {
            unsigned c = (unsigned) (xs[s] >> (2 * (k - 1 - j))) & 3u;
            f[j] = (double) c * 2.0 / 3.0 - 1.0;
        }
        double out = b2[0];
        for (uint32_t i = 0; i < h; i++) {
            double a = b1[i];
            for (uint32_t j = 0; j < k; j++)
                a += W1[(size_t) i * k + j] * f[j];
            pre[i] = a;
            if (a > 0)
                out += w2[i] * a;
        }
        double r = out - targets[s];
        loss += r * r;
        if (!grad)
            continue;
        double e = 2.0 * r / (double) n;   /* d loss / d out */
        double *gW1 = grad, *gb1 = grad + (size_t) h * k, *gw2 = gb1 + h, *gb2 = gw2 + h;
        gb2[0] += e;
        for (uint32_t i = 0; i < h; i++) {
            if (pre[i] <= 0)
                continue;
            gw2[i] += e * pre[i];
            double dp = e * w2[i];
            gb1[i] += dp;
            for (uint32_t j = 0; j < k; j++)
                gW1[(size_t) i * k + j] += dp * f[j];
        }
    }
    free(f);
    free(pre);
    return loss / (double) n;
}

void orc_mlp_forward(uint32_t k, uint32_t h, const double *params, const uint64_t *xs,
                     uint64_t n, double *out)
{
    const double *W1 = params, *b1 = params + (size_t) h * k, *w2 = b1 + h, *b2 = w2 + h;
    for (uint64_t s = 0; s < n; s++) {
        double y = b2[0];
        for (uint32_t i = 0; i < h; i++) {
            double a = b1[i];
            for (uint32_t j = 0; j < k; j++) {
                unsigned c = (unsigned) (xs[s] >> (2 * (k - 1 - j))) & 3u;
                a += W1[(size_t) i * k + j] * ((double) c * 2.0 / 3.0 - 1.0);
            }
            y += a > 0 ? w2[i] * a : 0.0;
        }
        out[s] = y;
    }
}

/* ---- first-triplet slot sizes (sufsort.c:386-450) ------------------------------ */

void orc_slot_sizes(const uint8_t *dtext, const uint64_t *off, uint64_t n, uint64_t size[5],
                    uint64_t offset[5], uint64_t *max_size)
{
    uint64_t freq[85] = { 0 };
    for (uint64_t r = 0; r < n; r++)
        for (uint64_t g = off[r]; g < off[r + 1]; g++) {
            const uint64_t rem = off[r + 1] - g;
            unsigned code = 1 + 21 * (dtext[g] - 1);
            if (rem >= 2) {
                code += 1 + 5 * (dtext[g + 1] - 1);
                if (rem >= 3)
                    code += 1 + (dtext[g + 2] - 1);
            }
            freq[code]++;
        }
    *max_size = 0;
    for (int k = 0; k < 5; k++) {
        size[k] = 0;
        for (int j = 0; j < 17; j++)
            size[k] += freq[17 * k + j];
        offset[k] = k ? offset[k - 1] + size[k - 1] : 0;
        if (size[k] > *max_size)
            *max_size = size[k];
    }
}

/* ---- reads import with duplicate removal (index.c:145-405, bloom.c) ---------- */

/* The hashes of bloom.c:18-86 over a NUL-terminated code string, in the index
 * width: wide = 0 reduces every step modulo 2^32. */
static uint64_t orc_hash(int f, const uint8_t *k, uint64_t size, int wide)
{
    const uint64_t mask = wide ? ~0ull : 0xFFFFFFFFull;
    uint64_t h = 0;
    for (; *k; k++) {
        const uint64_t c = *k;
        switch (f) {
        case 0: h ^= ((h << 5) + (h >> 2) + c) & mask; break;
        case 1: h = c + (h << 6) + (h << 16) - h; break;
        case 2: h = h + c; break;
        case 3: h ^= c; break;
        default: h = 33 * h + c; break;
        }
        h &= mask;
    }
    return h % size;
}

/* bloom_filter, bloom.c:91-135: 1 = seen (read or its reverse complement),
 * 0 = new, its own bits set. */
static int orc_bloom(uint8_t *a, uint64_t size, int nf, const uint8_t *s,
                     const uint8_t *rs, int wide)
{
    uint64_t hv[5];
    int miss = 0;
    for (int i = 0; i < nf; i++) {
        hv[i] = orc_hash(i, s, size, wide);
        if (!(a[hv[i] / 8] & (1u << (hv[i] % 8))))
            miss = 1;
    }
    if (!miss)
        return 1;
    for (int i = 0; i < nf; i++) {
        uint64_t r = orc_hash(i, rs, size, wide);
        if (!(a[r / 8] & (1u << (r % 8)))) {
            for (int j = 0; j < nf; j++)
                a[hv[j] / 8] |= (uint8_t) (1u << (hv[j] % 8));
            return 0;
        }
    }
    return 1;
}

void orc_dedup(const uint8_t *codes, const uint8_t *ok, uint64_t nrec, int read_len,
               uint64_t bloom_bits, int n_hashfunc, int index_bits, uint8_t *keep)
{
    const int L = read_len, wide = index_bits == 64;
    uint8_t *a = calloc((size_t) (bloom_bits / 8 + 2), 1);
    uint8_t *s = malloc((size_t) L + 1), *rs = malloc((size_t) L + 1);
    for (uint64_t r = 0; r < nrec; r++) {
        keep[r] = 0;
        if (!ok[r])
            continue;
        for (int i = 0; i < L; i++) {
            s[i] = codes[r * (uint64_t) L + i];
            rs[L - 1 - i] = (uint8_t) (5 - s[i]);
        }
        s[L] = rs[L] = 0;
        keep[r] = orc_bloom(a, bloom_bits, n_hashfunc, s, rs, wide) ? 0 : 1;
    }
    free(a);
    free(s);
    free(rs);
}

int orc_import_fasta(const char *path, int read_len, int bloom_m, int n_hashfunc,
                     int index_bits, uint8_t **reads, uint64_t *n, uint64_t *removed,
                     long *bad_line)
{
    char line[1000];
    FILE *fp = fopen(path, "rb");
    *reads = NULL;
    *n = *removed = 0;
    *bad_line = 0;
    if (!fp)
        return -2;
    uint64_t records = 0;
    while (fgets(line, 1000, fp))
        records += line[0] == '>';
    const int wide = index_bits == 64;
    const uint64_t bits = wide ? records * (uint64_t) bloom_m
                               : (uint64_t) (uint32_t) ((uint32_t) records * (uint32_t) bloom_m);
    uint8_t *a = calloc((size_t) (bits / 8 + 2), 1);
    uint8_t *s = malloc((size_t) read_len + 1), *rs = malloc((size_t) read_len + 1);
    uint64_t cap = 1024, kept = 0, dropped = 0;
    uint8_t *out = malloc(cap * (size_t) read_len);
    int rc = (a && s && rs && out) ? 0 : -5;
    rewind(fp);
    while (rc == 0 && fgets(line, 1000, fp)) {
        if (line[0] != '>' || !fgets(line, 1000, fp)) {
            rc = -3;
            break;
        }
        size_t len = strlen(line);
        if (len == 0 || len - 1 != (size_t) read_len) {
            rc = -3;
            break;
        }
        int with_n = 0;
        for (int i = 0; i < read_len && rc == 0; i++) {
            uint8_t v = 0;
            switch (line[i]) {
            case 'a': case 'A': v = 1; break;
            case 'c': case 'C': v = 2; break;
            case 'g': case 'G': v = 3; break;
            case 't': case 'T': v = 4; break;
            case 'n': case 'N': with_n = 1; break;
            default: rc = -3;
            }
            s[i] = v;
            rs[read_len - 1 - i] = (uint8_t) (v ? 5 - v : 0);
        }
        if (rc)
            break;
        s[read_len] = rs[read_len] = 0;
        if (with_n) {
            dropped++;
            continue;
        }
        if (bits == 0) {            /* h % 0: the reference faults */
            rc = -7;
            break;
        }
        if (orc_bloom(a, bits, n_hashfunc, s, rs, wide)) {
            dropped++;
            continue;
        }
        if (kept == cap) {
            uint8_t *nb = realloc(out, 2 * cap * (size_t) read_len);
            if (!nb) {
                rc = -5;
                break;
            }
            out = nb;
            cap *= 2;
        }
        memcpy(out + kept * (uint64_t) read_len, s, (size_t) read_len);
        kept++;
    }
    fclose(fp);
    free(a);
    free(s);
    free(rs);
    if (rc == -3)
        *bad_line = (long) (kept * 2 + 2);
    if (rc) {
        free(out);
        return rc;
    }
    *reads = out;
    *n = kept;
    *removed = dropped;
    return 0;
}


/* ---- work of the device's candidate search ----------------------------------- */

/* rank_c(i) = #{j < i : bwt[j] == c} from per-64-row counts and a scan. */
typedef struct {
    const uint8_t *bwt;
    uint64_t *blk;     /* 5 counts per 64-row block, before the block */
} orc_rank_t;

static uint64_t chain_rank(const orc_rank_t *R, unsigned c, uint64_t i)
{
    uint64_t b = i >> 6, k = R->blk[b * 5 + c];
    for (uint64_t j = b << 6; j < i; j++)
        k += R->bwt[j] == c;
    return k;
}

/* The code a chain prepends at depth o: n[L-o] (FB), its complement (FF),
 * n[o-1] (BF), its complement (BB). */
static unsigned chain_code(const uint8_t *s, uint32_t L, int ch, uint32_t o)
{
    const unsigned x = ch < 2 ? s[L - o] : s[o - 1];
    return (ch == 1 || ch == 3) ? 5u - x : x;
}

/* Rows the learned seed kernel's outward scan reads for a pattern with rows
 * [l, l + size) of an index of N rows sampled every st-th row (ns samples),
 * or 0 when it takes the binary searches instead (*scan = 0).  jl = the first
 * sample at or after row l, jh = the first at or after row l + size; with a
 * sample inside the interval the scan starts there and reads the members plus
 * one row past each end that exists above the previous sample; without one it
 * reads down from the sample's row to one row past the interval. */
static uint64_t seed_scan_rows(uint64_t l, uint64_t size, uint64_t N, uint64_t st, uint32_t vmax,
                               int *scan)
{
    const uint64_t ns = (N + st - 1) / st;
    uint64_t jl = (l + st - 1) / st, jh = (l + size + st - 1) / st;
    if (jl > ns)
        jl = ns;
    if (jh > ns)
        jh = ns;
    const int m1 = jl < ns && jl * st < l + size;
    const uint64_t bl = jl > 0 ? (jl - 1) * st + 1 : 0, r1 = jl < ns ? jl * st : N;
    *scan = vmax > 0 && (!m1 || (jh - jl - 1) * st + 1 <= vmax);
    if (!*scan)
        return 0;
    if (m1)
        return size + (l > bl) + (l + size < N);
    return (r1 - l) + (l > bl);
}

int orc_chain_work(const uint8_t *text, const uint64_t *off, uint64_t n, uint32_t seed_depth,
                   uint32_t t, uint32_t vmax, uint32_t stride, uint64_t out[3])
{
    out[0] = out[1] = out[2] = 0;
    const uint64_t N = off[n];
    uint64_t C[6], sfx[4], pfx[4];
    orc_counts(text, off, n, C, sfx, pfx);
    uint8_t *dt = malloc(N + 1), *bwt[2] = {malloc(N + 1), malloc(N + 1)};
    uint64_t *sa = malloc(8 * (N + 1)), *ri = malloc(8 * (n + 1));
    uint64_t *blk[2] = {malloc(8 * 5 * (N / 64 + 2)), malloc(8 * 5 * (N / 64 + 2))};
    int rc = 0;
    if (!dt || !bwt[0] || !bwt[1] || !sa || !ri || !blk[0] || !blk[1]) {
        rc = -5;
        goto done;
    }
    orc_rank_t R[2];
    uint64_t cf[2][5];
    for (int d = 0; d < 2; d++) {
        orc_dir_text(text, off, n, d, dt);
        if (orc_sa(dt, off, n, sa)) {
            rc = -5;
            goto done;
        }
        orc_bwt(dt, off, n, sa, bwt[d], ri);
        uint64_t k[5] = {0, 0, 0, 0, 0};
        for (uint64_t i = 0; i <= N; i++) {
            if ((i & 63) == 0)
                memcpy(blk[d] + (i >> 6) * 5, k, sizeof(k));
            if (i < N)
                k[bwt[d][i]]++;
        }
        R[d].bwt = bwt[d];
        R[d].blk = blk[d];
        /* rows of suffixes starting with code c: after all rows starting with a
         * smaller code (C[c]) and the one-code suffixes "c" that end a read of
         * this direction (they follow no row); in the order of their tails */
        for (unsigned c = 1; c <= 4; c++)
            cf[d][c] = C[c] + (d ? pfx[c - 1] : sfx[c - 1]);
        cf[d][0] = 0;
    }
    for (uint64_t r = 0; r < n; r++) {
        const uint8_t *s = text + off[r];
        const uint32_t L = (uint32_t) (off[r + 1] - off[r]);
        for (int ch = 0; ch < 4; ch++) {
            const int d = (ch == 0 || ch == 3) ? 0 : 1;
            const uint64_t minsize = (ch == 0 || ch == 2) ? 2 : 1;
            unsigned c = chain_code(s, L, ch, 1);
            uint64_t l = C[c], size = C[c + 1] - C[c];
            uint32_t o = 1;
            if (seed_depth >= 1 && seed_depth < L) {
                for (; o < seed_depth; o++) {
                    c = chain_code(s, L, ch, o + 1);
                    const uint64_t a = chain_rank(&R[d], c, l), b = chain_rank(&R[d], c, l + size);
                    l = cf[d][c] + a;
                    size = b - a;
                }
                int scan = 0;
                if (stride && o <= t)
                    out[2] += seed_scan_rows(l, size, N, stride, vmax, &scan);
                /* a scanned interval of minsize..vmax rows is finished there */
                if (scan && size >= minsize && size <= vmax)
                    continue;
                /* otherwise size < minsize or size > vmax: the walk below goes on */
            }
            while (o < L && size >= minsize) {
                if (size <= vmax && o <= t) {
                    out[2] += size;
                    break;
                }
                out[0] += 2;
                out[1] += (l >> 7) == ((l + size) >> 7) ? 1 : 2;
                if (o + 1 >= L)
                    break;
                c = chain_code(s, L, ch, o + 1);
                const uint64_t a = chain_rank(&R[d], c, l), b = chain_rank(&R[d], c, l + size);
                l = cf[d][c] + a;
                size = b - a;
                o++;
            }
        }
    }
done:
    free(dt);
    free(bwt[0]);
    free(bwt[1]);
    free(sa);
    free(ri);
    free(blk[0]);
    free(blk[1]);
    return rc;
}
