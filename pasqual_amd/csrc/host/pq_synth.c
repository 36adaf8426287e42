/*
 * pq_synth.c -- deterministic synthetic genomes and reads.
 *
 * PRNG: xoshiro256** seeded through splitmix64.  The genome (and its repeat
 * copies) comes from one stream seeded by cfg->seed; read r comes from its own
 * stream seeded by splitmix64 of (seed + (r+1) * golden gamma), so reads can be
 * produced in parallel and any read can be regenerated alone.  Per read the draw
 * order is: length (only if len_min != len_max), start, strand, then one error
 * draw per base (plus one substitute draw for each error).
 *
 * Reads are single-end, start uniform in [0, G - len], reverse-complemented
 * with probability 1/2, with i.i.d. substitutions to one of the three other
 * bases.  Output codes are 1..4 for A,C,G,T (index.c:280-321).
 */
#include "../../../include/pasqual_amd.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++)
        r->s[i] = splitmix64(&x);
}

static uint64_t rotl(uint64_t x, int k)
{
    return (x << k) | (x >> (64 - k));
}

static uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

/* uniform integer in [0, m) by multiply-shift */
static uint64_t rng_below(rng_t *r, uint64_t m)
{
    return (uint64_t) (((__uint128_t) rng_next(r) * m) >> 64);
}

static uint64_t prob_threshold(double p)
{
    if (p <= 0.0)
        return 0;
    if (p >= 1.0)
        return UINT64_MAX;
    return (uint64_t) ldexp(p, 64);
}

static uint8_t mutate(rng_t *r, uint8_t c)
{
    return (uint8_t) (((c - 1) + 1 + rng_below(r, 3)) % 4 + 1);
}

int pq_synth_preset(int config, pq_synth_cfg_t *c)
{
    memset(c, 0, sizeof(*c));
    c->seed = 0x5EED0000ULL + (uint64_t) config;
    switch (config) {
    case 1:                        /* 1 Mbp uniform, 50 bp, 20x, error-free */
        c->genome_len = 1000000;
        c->len_min = c->len_max = 50;
        c->coverage = 20;
        c->min_overlap = 31;
        break;
    case 2:                        /* 5 Mbp + 100 copies of ten 1 kbp units, 1% div */
        c->genome_len = 5000000;
        c->n_units = 10;
        c->unit_len = 1000;
        c->n_copies = 100;
        c->copy_div = 0.01;
        c->len_min = c->len_max = 100;
        c->coverage = 40;
        c->sub_rate = 0.005;
        c->min_overlap = 63;
        break;
    case 3:                        /* 50 Mbp, lengths U{36..250}, 30x, 0.3% */
    case 4:
        c->genome_len = 50000000;
        c->len_min = 36;
        c->len_max = 250;
        c->coverage = 30;
        c->sub_rate = 0.003;
        c->min_overlap = 31;
        break;
    case 5:                        /* 200 Mbp + 5% repeat families, 150 bp, 30x */
        c->genome_len = 200000000;
        c->family_frac = 0.05;
        c->fam_len_min = 1000;
        c->fam_len_max = 10000;
        c->fam_div_min = 0.005;
        c->fam_div_max = 0.03;
        c->fam_copies_min = 2;
        c->fam_copies_max = 50;
        c->len_min = c->len_max = 150;
        c->coverage = 30;
        c->sub_rate = 0.003;
        c->min_overlap = 63;
        break;
    default:
        return -1;
    }
    return 0;
}

uint64_t pq_synth_num_reads(const pq_synth_cfg_t *c)
{
    if (c->n_reads)
        return c->n_reads;
    double mean = 0.5 * ((double) c->len_min + (double) c->len_max);
    return (uint64_t) ceil(c->coverage * (double) c->genome_len / mean);
}

int pq_synth_genome(const pq_synth_cfg_t *c, uint8_t *g)
{
    uint64_t G = c->genome_len;
    rng_t r;
    rng_seed(&r, c->seed ^ 0xA5A5A5A5DEADBEEFULL);
    for (uint64_t i = 0; i < G; i += 32) {
        uint64_t x = rng_next(&r);
        for (uint64_t j = 0; j < 32 && i + j < G; j++)
            g[i + j] = (uint8_t) (((x >> (2 * j)) & 3) + 1);
    }
    /* interspersed copies of repeat units in disjoint unit-aligned slots */
    if (c->n_units && c->unit_len && c->n_copies) {
        uint64_t U = c->unit_len, slots = G / U;
        if (slots < c->n_copies)
            return -1;
        uint8_t *units = malloc((size_t) (c->n_units * U));
        uint64_t *perm = malloc(sizeof(uint64_t) * slots);
        if (!units || !perm) {
            free(units);
            free(perm);
            return -1;
        }
        for (uint64_t i = 0; i < c->n_units * U; i++)
            units[i] = (uint8_t) (rng_below(&r, 4) + 1);
        for (uint64_t i = 0; i < slots; i++)
            perm[i] = i;
        for (uint64_t i = 0; i < c->n_copies; i++) {
            uint64_t j = i + rng_below(&r, slots - i), tmp = perm[i];
            perm[i] = perm[j];
            perm[j] = tmp;
        }
        uint64_t thr = prob_threshold(c->copy_div);
        for (uint64_t k = 0; k < c->n_copies; k++) {
            const uint8_t *u = units + (k % c->n_units) * U;
            uint8_t *dst = g + perm[k] * U;
            for (uint64_t i = 0; i < U; i++)
                dst[i] = rng_next(&r) < thr ? mutate(&r, u[i]) : u[i];
        }
        free(units);
        free(perm);
    }
    /* repeat families: random consensus, copies at random places */
    if (c->family_frac > 0 && c->fam_len_max >= c->fam_len_min && c->fam_len_min) {
        uint64_t target = (uint64_t) (c->family_frac * (double) G), placed = 0;
        uint8_t *cons = malloc(c->fam_len_max);
        if (!cons)
            return -1;
        while (placed < target) {
            uint64_t len = c->fam_len_min + rng_below(&r, c->fam_len_max - c->fam_len_min + 1);
            double div = c->fam_div_min + (c->fam_div_max - c->fam_div_min) *
                ((double) (rng_next(&r) >> 11) * 0x1.0p-53);
            uint64_t copies = c->fam_copies_min +
                rng_below(&r, c->fam_copies_max - c->fam_copies_min + 1);
            uint64_t thr = prob_threshold(div);
            if (len >= G)
                break;
            for (uint64_t i = 0; i < len; i++)
                cons[i] = (uint8_t) (rng_below(&r, 4) + 1);
            for (uint64_t k = 0; k < copies && placed < target; k++) {
                uint64_t pos = rng_below(&r, G - len + 1);
                for (uint64_t i = 0; i < len; i++)
                    g[pos + i] = rng_next(&r) < thr ? mutate(&r, cons[i]) : cons[i];
                placed += len;
            }
        }
        free(cons);
    }
    return 0;
}

static void read_rng(const pq_synth_cfg_t *c, uint64_t rid, rng_t *r)
{
    rng_seed(r, c->seed + (rid + 1) * 0xD1B54A32D192ED03ULL);
}

static uint32_t read_len_of(const pq_synth_cfg_t *c, rng_t *r)
{
    if (c->len_max == c->len_min)
        return c->len_min;
    return c->len_min + (uint32_t) rng_below(r, c->len_max - c->len_min + 1);
}

int pq_synth_lengths(const pq_synth_cfg_t *c, uint64_t *off)
{
    uint64_t n = pq_synth_num_reads(c);
    off[0] = 0;
    for (uint64_t i = 0; i < n; i++) {
        rng_t r;
        read_rng(c, i, &r);
        uint32_t L = read_len_of(c, &r);
        if (L > c->genome_len)
            return -1;
        off[i + 1] = off[i] + L;
    }
    return 0;
}

static void make_read(const pq_synth_cfg_t *c, const uint8_t *g, uint64_t rid,
                      uint8_t *out)
{
    rng_t r;
    read_rng(c, rid, &r);
    uint32_t L = read_len_of(c, &r);
    uint64_t s = rng_below(&r, c->genome_len - L + 1);
    int rc = (int) (rng_next(&r) >> 63);
    uint64_t thr = prob_threshold(c->sub_rate);
    for (uint32_t i = 0; i < L; i++) {
        uint8_t b = rc ? (uint8_t) (5 - g[s + L - 1 - i]) : g[s + i];
        if (thr && rng_next(&r) < thr)
            b = mutate(&r, b);
        out[i] = b;
    }
}

int pq_synth_reads(const pq_synth_cfg_t *c, const uint8_t *genome,
                   const uint64_t *off, uint8_t *codes)
{
    int64_t n = (int64_t) pq_synth_num_reads(c);
#pragma omp parallel for schedule(static, 4096)
    for (int64_t i = 0; i < n; i++)
        make_read(c, genome, (uint64_t) i, codes + off[i]);
    return 0;
}

int pq_synth_fasta(const pq_synth_cfg_t *c, const char *path)
{
    uint64_t n = pq_synth_num_reads(c);
    uint8_t *g = malloc(c->genome_len ? c->genome_len : 1);
    uint64_t *off = malloc(sizeof(uint64_t) * (n + 1));
    FILE *f = fopen(path, "w");
    int rc = -1;
    if (!g || !off || !f)
        goto out;
    if (pq_synth_genome(c, g) || pq_synth_lengths(c, off))
        goto out;
    static const char acgt[5] = { 'N', 'A', 'C', 'G', 'T' };
    char *line = malloc(c->len_max + 2);
    uint8_t *buf = malloc(c->len_max + 1);
    if (!line || !buf) {
        free(line);
        free(buf);
        goto out;
    }
    for (uint64_t i = 0; i < n; i++) {
        uint32_t L = (uint32_t) (off[i + 1] - off[i]);
        make_read(c, g, i, buf);
        for (uint32_t j = 0; j < L; j++)
            line[j] = acgt[buf[j]];
        line[L] = '\n';
        fprintf(f, ">r%llu\n", (unsigned long long) i);
        fwrite(line, 1, L + 1, f);
    }
    free(line);
    free(buf);
    rc = 0;
out:
    if (f && fclose(f))
        rc = -1;
    free(g);
    free(off);
    return rc;
}
