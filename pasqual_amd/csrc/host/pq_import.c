/*
 * pq_import.c -- reads-file import with duplicate removal.
 *
 * Follows import_singlereads (index.c:145-405): a two-line FASTA read through
 * a 1000-byte fgets buffer (index.c:218; a longer line arrives as several
 * pieces of at most 999 bytes), the last byte of each sequence line dropped as
 * its newline (index.c:253), exact read length required, A/C/G/T (either case)
 * mapped to 1..4, reads containing N dropped, and a Bloom filter of
 * bloom_m * #records bits (records = pieces starting with '>', index.c:195-206)
 * probed with the read and its reverse complement (bloom.c:91-135); kept reads
 * get ids 0..n-1 in input order.
 *
 * The reference does all of this in one sequential loop.  Here the file is
 * read whole, cut into fgets pieces and records by all host threads, the
 * order-dependent filter is resolved on the device (pq_dedup_reads, exact),
 * and the kept reads are gathered by all threads again.
 */
#include "../../../include/pasqual_amd.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>

#define PIECE_MAX 999     /* bytes fgets(line, 1000, fp) returns at most */

static double now_ms(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

/* The whole file in one buffer (size from fstat for a regular file, else read
 * until end of file). */
static int read_whole(const char *path, char **buf, uint64_t *size)
{
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return PQ_ERR_IO;
    struct stat st;
    uint64_t cap = 1 << 20;
    if (fstat(fileno(fp), &st) == 0 && S_ISREG(st.st_mode))
        cap = (uint64_t) st.st_size + 1;
    char *b = malloc(cap);
    uint64_t len = 0;
    int rc = b ? PQ_OK : PQ_ERR_NOMEM;
    while (rc == PQ_OK) {
        if (len == cap) {
            char *nb = realloc(b, cap * 2);
            if (!nb) {
                rc = PQ_ERR_NOMEM;
                break;
            }
            b = nb;
            cap *= 2;
        }
        size_t got = fread(b + len, 1, cap - len, fp);
        len += got;
        if (got == 0) {
            if (ferror(fp))
                rc = PQ_ERR_IO;
            break;
        }
    }
    fclose(fp);
    if (rc != PQ_OK) {
        free(b);
        return rc;
    }
    *buf = b;
    *size = len;
    return PQ_OK;
}

/* Start offsets of the pieces fgets would return, in file order; *np pieces,
 * (*pieces)[*np] = size.  A line (up to and including its newline, or to the
 * end of the file) of len bytes is ceil(len / 999) pieces. */
static int cut_pieces(const char *b, uint64_t size, uint64_t **pieces, uint64_t *np)
{
    const int T = omp_get_max_threads();
    uint64_t *cnt = calloc((size_t) T + 1, sizeof(uint64_t));
    if (!cnt)
        return PQ_ERR_NOMEM;
    const uint64_t seg = (size + T - 1) / T;
    /* pieces of the lines that start in each thread's byte range */
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = seg * t < size ? seg * t : size;
        const uint64_t e = seg * (t + 1) < size ? seg * (t + 1) : size;
        uint64_t c = 0;
        for (uint64_t s = a; s < e;) {
            if (s > 0 && b[s - 1] != '\n') {          /* not a line start: skip to one */
                const char *nl = memchr(b + s, '\n', e - s);
                if (!nl)
                    break;
                s = (uint64_t) (nl - b) + 1;
                continue;
            }
            const char *nl = memchr(b + s, '\n', size - s);
            const uint64_t end = nl ? (uint64_t) (nl - b) + 1 : size;
            c += (end - s + PIECE_MAX - 1) / PIECE_MAX;
            s = end;
        }
        cnt[t + 1] = c;
    }
    for (int t = 0; t < T; t++)
        cnt[t + 1] += cnt[t];
    const uint64_t n = cnt[T];
    uint64_t *p = malloc(sizeof(uint64_t) * (n + 1));
    if (!p) {
        free(cnt);
        return PQ_ERR_NOMEM;
    }
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = seg * t < size ? seg * t : size;
        const uint64_t e = seg * (t + 1) < size ? seg * (t + 1) : size;
        uint64_t k = cnt[t];
        for (uint64_t s = a; s < e;) {
            if (s > 0 && b[s - 1] != '\n') {
                const char *nl = memchr(b + s, '\n', e - s);
                if (!nl)
                    break;
                s = (uint64_t) (nl - b) + 1;
                continue;
            }
            const char *nl = memchr(b + s, '\n', size - s);
            const uint64_t end = nl ? (uint64_t) (nl - b) + 1 : size;
            for (uint64_t q = s; q < end; q += PIECE_MAX)
                p[k++] = q;
            s = end;
        }
    }
    p[n] = size;
    free(cnt);
    *pieces = p;
    *np = n;
    return PQ_OK;
}

enum { REC_OK = 0, REC_N = 1, REC_BAD = 2 };

/* Record r = pieces 2r (header) and 2r+1 (sequence): its codes and status,
 * index.c:220-321. */
static int parse_record(const char *b, const uint64_t *pc, uint64_t r, int L, uint8_t *codes)
{
    if (b[pc[2 * r]] != '>')
        return REC_BAD;
    const uint64_t s = pc[2 * r + 1], e = pc[2 * r + 2];
    const char *z = memchr(b + s, '\0', e - s);     /* strlen stops at a NUL */
    const uint64_t slen = z ? (uint64_t) (z - (b + s)) : e - s;
    if (slen == 0 || slen - 1 != (uint64_t) L)
        return REC_BAD;
    int st = REC_OK;
    for (int i = 0; i < L; i++) {
        uint8_t v;
        switch (b[s + i]) {
        case 'a': case 'A': v = 1; break;
        case 'c': case 'C': v = 2; break;
        case 'g': case 'G': v = 3; break;
        case 't': case 'T': v = 4; break;
        case 'n': case 'N': v = 0; st = REC_N; break;
        default: return REC_BAD;
        }
        codes[i] = v;
    }
    return st;
}

int pq_import_fasta(pq_ctx_t *ctx, const char *path, int read_len, int bloom_m, int n_hashfunc,
                    int index_bits, pq_import_t *out)
{
    const double t0 = now_ms();
    if (!out)
        return PQ_ERR_ARG;
    memset(out, 0, sizeof(*out));
    if (!ctx || !path || read_len <= 0 || n_hashfunc < 1 || n_hashfunc > 5 ||
        (index_bits != 32 && index_bits != 64))
        return PQ_ERR_ARG;
    const uint64_t L = (uint64_t) read_len;
    char *b = NULL;
    uint64_t size = 0, *pc = NULL, np = 0;
    int rc = read_whole(path, &b, &size);
    if (rc)
        return rc;
    if ((rc = cut_pieces(b, size, &pc, &np))) {
        free(b);
        return rc;
    }
    /* records for the filter size: pieces that start with '>' (index.c:195-206) */
    uint64_t records = 0;
#pragma omp parallel for reduction(+ : records)
    for (long long i = 0; i < (long long) np; i++)
        records += b[pc[i]] == '>';
    const uint64_t bsize = index_bits == 32
                               ? (uint64_t) (uint32_t) ((uint32_t) records * (uint32_t) bloom_m)
                               : records * (uint64_t) bloom_m;
    out->bloom_bits = bsize;

    /* every record up to the first malformed one (the reference stops there;
     * an odd last piece is a header without its sequence line) */
    const uint64_t npairs = np / 2;
    uint8_t *codes = malloc(npairs * L + 1), *ok = malloc(npairs + 1), *keep = malloc(npairs + 1);
    if (!codes || !ok || !keep) {
        rc = PQ_ERR_NOMEM;
        goto done;
    }
    uint64_t first_bad = np & 1 ? npairs : UINT64_MAX;
#pragma omp parallel for schedule(static, 4096) reduction(min : first_bad)
    for (long long r = 0; r < (long long) npairs; r++) {
        const int st = parse_record(b, pc, (uint64_t) r, read_len, codes + (uint64_t) r * L);
        ok[r] = st == REC_OK;
        if (st == REC_BAD && (uint64_t) r < first_bad)
            first_bad = (uint64_t) r;
    }
    const uint64_t P = first_bad < npairs ? first_bad : npairs;
    uint32_t passes = 0;
    if ((rc = pq_dedup_reads(ctx, codes, ok, P, read_len, bsize, n_hashfunc, index_bits, keep,
                             &passes)))
        goto done;

    /* kept reads in input order, gathered by thread ranges */
    const int T = omp_get_max_threads();
    uint64_t *kc = calloc((size_t) T + 1, sizeof(uint64_t));
    uint64_t (*cnt)[14] = calloc((size_t) T, sizeof(*cnt));   /* nt[4] prefix[4] suffix[4] */
    if (!kc || !cnt) {
        free(kc);
        free(cnt);
        rc = PQ_ERR_NOMEM;
        goto done;
    }
    const uint64_t seg = (P + T - 1) / T;
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = seg * t < P ? seg * t : P, e = seg * (t + 1) < P ? seg * (t + 1) : P;
        uint64_t c = 0;
        for (uint64_t r = a; r < e; r++)
            c += keep[r];
        kc[t + 1] = c;
    }
    for (int t = 0; t < T; t++)
        kc[t + 1] += kc[t];
    const uint64_t n = kc[T];
    if (first_bad != UINT64_MAX) {
        out->bad_line = (long) (n * 2 + 2);      /* index.c:223 */
        rc = PQ_ERR_FORMAT;
        free(kc);
        free(cnt);
        goto done;
    }
    uint8_t *reads = malloc(n * L + 1);
    if (!reads) {
        free(kc);
        free(cnt);
        rc = PQ_ERR_NOMEM;
        goto done;
    }
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = seg * t < P ? seg * t : P, e = seg * (t + 1) < P ? seg * (t + 1) : P;
        uint64_t k = kc[t];
        uint64_t *c = cnt[t];
        for (uint64_t r = a; r < e; r++) {
            if (!keep[r])
                continue;
            const uint8_t *s = codes + r * L;
            memcpy(reads + k * L, s, L);
            k++;
            c[4 + s[0] - 1]++;
            c[8 + s[L - 1] - 1]++;
            for (uint64_t i = 0; i < L; i++)
                c[s[i] - 1]++;
        }
    }
    uint64_t nt[4] = { 0, 0, 0, 0 };
    for (int t = 0; t < T; t++)
        for (int c = 0; c < 4; c++) {
            nt[c] += cnt[t][c];
            out->prefix_count[c] += cnt[t][4 + c];
            out->suffix_count[c] += cnt[t][8 + c];
        }
    free(kc);
    free(cnt);
    out->reads = reads;
    out->n = n;
    out->removed = P - n;
    out->C[0] = 0;
    out->C[1] = 0;
    out->C[2] = nt[0];
    out->C[3] = nt[0] + nt[1];
    out->C[4] = nt[0] + nt[1] + nt[2];
    out->C[5] = nt[0] + nt[1] + nt[2] + nt[3];
done:
    free(codes);
    free(ok);
    free(keep);
    free(pc);
    free(b);
    out->ms = now_ms() - t0;
    return rc;
}
