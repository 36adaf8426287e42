/*
 * pq_import.c -- reads-file import with duplicate removal.
 *
 * Host-side, sequential by nature (the duplicate filter depends on insertion
 * order).  Follows import_singlereads (index.c:145-405): a two-line FASTA read
 * with a 1000-byte line buffer, the last byte of each sequence line dropped as
 * its newline (index.c:253), exact read length required, A/C/G/T (either case)
 * mapped to 1..4, reads containing N dropped, and a Bloom filter of
 * bloom_m * #records bits probed with the read and its reverse complement
 * (bloom.c:91-135).  The hash functions are the five string hashes of
 * bloom.c:18-86, evaluated in the index width (INDEX_TYPE_T), which decides
 * which reads survive; kept reads get ids 0..n-1 in input order.
 */
#include "pq_host.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#define LINE_BUF 1000

/* ---- hashes over a NUL-terminated code string, in 32- or 64-bit arithmetic ---- */

#define DEF_HASHES(T, S)                                                      \
    static T h0_##S(const char *k, T size)                                    \
    {                                                                         \
        T h = 0;                                                              \
        while (*k)                                                            \
            h ^= (T) ((h << 5) + (h >> 2) + (unsigned char) *k++);            \
        return h % size;                                                      \
    }                                                                         \
    static T h1_##S(const char *k, T size)                                    \
    {                                                                         \
        T h = 0;                                                              \
        while (*k)                                                            \
            h = (T) ((unsigned char) *k++ + (h << 6) + (h << 16) - h);        \
        return h % size;                                                      \
    }                                                                         \
    static T h2_##S(const char *k, T size)                                    \
    {                                                                         \
        T h = 0;                                                              \
        while (*k)                                                            \
            h = (T) (h + (T) (signed char) *k++);                             \
        return h % size;                                                      \
    }                                                                         \
    static T h3_##S(const char *k, T size)                                    \
    {                                                                         \
        T h = 0;                                                              \
        while (*k)                                                            \
            h ^= (T) (signed char) *k++;                                      \
        return h % size;                                                      \
    }                                                                         \
    static T h4_##S(const char *k, T size)                                    \
    {                                                                         \
        T h = 0;                                                              \
        while (*k)                                                            \
            h = (T) (33 * h + (T) (signed char) *k++);                        \
        return h % size;                                                      \
    }                                                                         \
    static T hash_##S(int f, const char *k, T size)                           \
    {                                                                         \
        switch (f) {                                                          \
        case 0: return h0_##S(k, size);                                       \
        case 1: return h1_##S(k, size);                                       \
        case 2: return h2_##S(k, size);                                       \
        case 3: return h3_##S(k, size);                                       \
        default: return h4_##S(k, size);                                      \
        }                                                                     \
    }                                                                         \
    /* 1 = seen before (read or its reverse complement), 0 = new (bits set) */\
    static int seen_##S(unsigned char *a, T size, int nf, const char *s,      \
                        const char *rs)                                       \
    {                                                                         \
        T hv[5];                                                              \
        int miss = 0;                                                         \
        for (int i = 0; i < nf; i++) {                                        \
            hv[i] = hash_##S(i, s, size);                                     \
            if (!(a[hv[i] / 8] & (1 << (hv[i] % 8))))                         \
                miss = 1;                                                     \
        }                                                                     \
        if (!miss)                                                            \
            return 1;                                                         \
        for (int i = 0; i < nf; i++) {                                        \
            T r = hash_##S(i, rs, size);                                      \
            if (!(a[r / 8] & (1 << (r % 8)))) {                               \
                for (int j = 0; j < nf; j++)                                  \
                    a[hv[j] / 8] |= (unsigned char) (1 << (hv[j] % 8));       \
                return 0;                                                     \
            }                                                                 \
        }                                                                     \
        return 1;                                                             \
    }

DEF_HASHES(uint32_t, 32)
DEF_HASHES(uint64_t, 64)

int pq_import_fasta(const char *path, int read_len, int bloom_m, int n_hashfunc,
                    int index_bits, pq_import_t *out)
{
    struct timeval tv1, tv2;
    gettimeofday(&tv1, NULL);
    memset(out, 0, sizeof(*out));
    if (read_len <= 0 || n_hashfunc < 1 || n_hashfunc > 5 ||
        (index_bits != 32 && index_bits != 64))
        return PQ_ERR_ARG;
    FILE *fp = fopen(path, "r");
    if (!fp)
        return PQ_ERR_IO;
    char line[LINE_BUF];
    uint64_t records = 0;
    while (fgets(line, LINE_BUF, fp))
        if (line[0] == '>')
            records++;

    /* filter size in the index width (index.c:206) */
    uint64_t bsize = index_bits == 32 ? (uint64_t) (uint32_t) ((uint32_t) records * (uint32_t) bloom_m)
                                      : records * (uint64_t) bloom_m;
    unsigned char *bloom = calloc((size_t) ((bsize + 7) / 8) + 1, 1);
    char *rline = malloc((size_t) read_len + 1);
    uint64_t cap = 1 << 20;
    uint8_t *reads = malloc(cap * (size_t) read_len);
    if (!bloom || !rline || !reads) {
        free(bloom);
        free(rline);
        free(reads);
        fclose(fp);
        return PQ_ERR_NOMEM;
    }
    uint64_t n = 0, removed = 0, cnt[4] = { 0, 0, 0, 0 };
    int rc = PQ_OK;
    fseek(fp, 0, SEEK_SET);
    while (fgets(line, LINE_BUF, fp)) {
        if (line[0] != '>') {
            rc = PQ_ERR_FORMAT;
            break;
        }
        if (!fgets(line, LINE_BUF, fp)) {
            rc = PQ_ERR_FORMAT;
            break;
        }
        size_t sl = strlen(line);
        line[sl - 1] = '\0';      /* drops the newline (or the last code) */
        if (strlen(line) != (size_t) read_len) {
            rc = PQ_ERR_FORMAT;
            break;
        }
        int with_n = 0;
        for (int i = 0; i < read_len && rc == PQ_OK; i++) {
            char c = line[i], v = 0, rv = 0;
            switch (c) {
            case 'a': case 'A': v = 1; rv = 4; break;
            case 'c': case 'C': v = 2; rv = 3; break;
            case 'g': case 'G': v = 3; rv = 2; break;
            case 't': case 'T': v = 4; rv = 1; break;
            case 'n': case 'N': with_n = 1; v = c; break;
            default: rc = PQ_ERR_FORMAT;
            }
            line[i] = v;
            rline[read_len - 1 - i] = rv;
        }
        if (rc != PQ_OK)
            break;
        rline[read_len] = '\0';
        if (with_n) {
            removed++;
            continue;
        }
        int dup = index_bits == 32
            ? seen_32(bloom, (uint32_t) bsize, n_hashfunc, line, rline)
            : seen_64(bloom, bsize, n_hashfunc, line, rline);
        if (dup) {
            removed++;
            continue;
        }
        out->prefix_count[line[0] - 1]++;
        out->suffix_count[line[read_len - 1] - 1]++;
        for (int i = 0; i < read_len; i++)
            cnt[line[i] - 1]++;
        if (n == cap) {
            cap *= 2;
            uint8_t *nr = realloc(reads, cap * (size_t) read_len);
            if (!nr) {
                rc = PQ_ERR_NOMEM;
                break;
            }
            reads = nr;
        }
        memcpy(reads + n * (uint64_t) read_len, line, (size_t) read_len);
        n++;
    }
    if (rc == PQ_ERR_FORMAT)
        out->bad_line = (long) (n * 2 + 2);   /* as reported by index.c:223 */
    fclose(fp);
    free(bloom);
    free(rline);
    if (rc != PQ_OK) {
        free(reads);
        return rc;
    }
    out->reads = reads;
    out->n = n;
    out->removed = removed;
    out->C[0] = 0;
    out->C[1] = 0;
    out->C[2] = cnt[0];
    out->C[3] = cnt[0] + cnt[1];
    out->C[4] = cnt[0] + cnt[1] + cnt[2];
    out->C[5] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
    gettimeofday(&tv2, NULL);
    out->ms = (tv2.tv_sec - tv1.tv_sec) * 1000.0 + (tv2.tv_usec - tv1.tv_usec) / 1000.0;
    return PQ_OK;
}
