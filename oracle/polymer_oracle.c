/*
 * polymer_oracle.c -- CPU restatement of Polymer's six applications and of
 * the synthetic-input generators.  TEST INFRASTRUCTURE ONLY (the checker):
 * tests/, the smoke check and the benchmark's CPU-baseline leg use it; the
 * GPU library never links it.
 *
 * Built with -ffp-contract=off so that float expressions round exactly as
 * written (the GPU side uses __fmul_rn/__fadd_rn for the same values).
 */
#include "polymer_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* hashing / synthetic inputs                                          */
/* ------------------------------------------------------------------ */

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
uint64_t or_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_rmat_thresholds(double a, double b, double c, uint32_t thr[3]) {
    double t0 = floor(a * 4294967296.0);
    double t1 = floor((a + b) * 4294967296.0);
    double t2 = floor((a + b + c) * 4294967296.0);
    thr[0] = t0 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t0;
    thr[1] = t1 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t1;
    thr[2] = t2 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t2;
}

/* Edge k of the recursive-matrix sequence: one 32-bit uniform per level,
 * quadrant a=(0,0) b=(0,1) c=(1,0) d=(1,1), most significant bit first. */
void or_rmat_pair(uint64_t seed, int scale, const uint32_t thr[3], uint64_t k,
                  uint32_t *src, uint32_t *dst) {
    uint64_t base = or_mix64(seed ^ 0x5851F42D4C957F2Dull);
    uint32_t s = 0, d = 0;
    for (int i = 0; i < scale; i++) {
        uint32_t r = (uint32_t)(or_mix64(base + k * 32ull + (uint64_t)i) >> 32);
        uint32_t q = r < thr[0] ? 0u : (r < thr[1] ? 1u : (r < thr[2] ? 2u : 3u));
        s = (s << 1) | (q >> 1);
        d = (d << 1) | (q & 1u);
    }
    *src = s;
    *dst = d;
}

/* stable LSD radix sort of (key, ctr) by key, 16-bit digits */
static void radix_sort_kv(uint64_t *key, uint32_t *val, int64_t cnt, int key_bits) {
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cnt);
    uint32_t *v2 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cnt);
    int64_t *hist = (int64_t *)malloc(sizeof(int64_t) * 65536);
    for (int sh = 0; sh < key_bits; sh += 16) {
        memset(hist, 0, sizeof(int64_t) * 65536);
        for (int64_t i = 0; i < cnt; i++) hist[(key[i] >> sh) & 0xFFFF]++;
        int64_t acc = 0;
        for (int b = 0; b < 65536; b++) { int64_t h = hist[b]; hist[b] = acc; acc += h; }
        for (int64_t i = 0; i < cnt; i++) {
            int64_t p = hist[(key[i] >> sh) & 0xFFFF]++;
            k2[p] = key[i];
            v2[p] = val[i];
        }
        memcpy(key, k2, sizeof(uint64_t) * (size_t)cnt);
        memcpy(val, v2, sizeof(uint32_t) * (size_t)cnt);
    }
    free(k2); free(v2); free(hist);
}

int64_t or_rmat_edges(int scale, int64_t m, double a, double b, double c,
                      uint64_t seed, int32_t *src, int32_t *dst) {
    if (scale < 1 || scale > 30 || m < 0) return -1;
    uint32_t thr[3];
    or_rmat_thresholds(a, b, c, thr);
    int64_t C = m + m / 8 + 1024;
    for (;;) {
        if (C > 0xFFFFFFFFll) return -2;
        uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)C);
        uint32_t *ctr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)C);
        if (!key || !ctr) { free(key); free(ctr); return -3; }
        int64_t cnt = 0;
        for (int64_t k = 0; k < C; k++) {
            uint32_t s, d;
            or_rmat_pair(seed, scale, thr, (uint64_t)k, &s, &d);
            if (s == d) continue;
            key[cnt] = ((uint64_t)s << 32) | d;
            ctr[cnt] = (uint32_t)k;
            cnt++;
        }
        /* key bits: 32+scale (src in high word) -> sort low scale bits and
         * the src bits; sorting 64 bits is simplest and still exact. */
        radix_sort_kv(key, ctr, cnt, 64);
        /* keep first occurrence (minimal counter, stable sort) */
        uint8_t *first = (uint8_t *)calloc((size_t)C, 1);
        int64_t uniq = 0;
        for (int64_t i = 0; i < cnt; i++) {
            if (i == 0 || key[i] != key[i - 1]) { first[ctr[i]] = 1; uniq++; }
        }
        if (uniq >= m) {
            /* cutoff = counter of the m-th first occurrence in counter order */
            int64_t seen = 0, cutoff = -1;
            for (int64_t k = 0; k < C && seen < m; k++) {
                if (first[k]) { seen++; if (seen == m) cutoff = k; }
            }
            int64_t out = 0;
            for (int64_t i = 0; i < cnt && m > 0; i++) {
                if (i == 0 || key[i] != key[i - 1]) {
                    if ((int64_t)ctr[i] <= cutoff) {
                        src[out] = (int32_t)(key[i] >> 32);
                        dst[out] = (int32_t)(key[i] & 0xFFFFFFFFu);
                        out++;
                    }
                }
            }
            free(first); free(key); free(ctr);
            return out;
        }
        free(first); free(key); free(ctr);
        C = C + C / 2 + 1024;
    }
}

void or_csr_from_sorted(int64_t n, int64_t m, const int32_t *src,
                        const int32_t *dst, int64_t *off, int32_t *adj) {
    for (int64_t i = 0; i <= n; i++) off[i] = 0;
    for (int64_t e = 0; e < m; e++) off[src[e] + 1]++;
    for (int64_t i = 0; i < n; i++) off[i + 1] += off[i];
    for (int64_t e = 0; e < m; e++) adj[e] = dst[e];
}

void or_transpose(int64_t n, int64_t m, const int64_t *off, const int32_t *adj,
                  int64_t *toff, int32_t *tadj, int64_t *eperm) {
    for (int64_t i = 0; i <= n; i++) toff[i] = 0;
    for (int64_t e = 0; e < m; e++) toff[adj[e] + 1]++;
    for (int64_t i = 0; i < n; i++) toff[i + 1] += toff[i];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    memcpy(fill, toff, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t s = 0; s < n; s++) {
        for (int64_t e = off[s]; e < off[s + 1]; e++) {
            int64_t p = fill[adj[e]]++;
            tadj[p] = (int32_t)s;
            if (eperm) eperm[p] = e;
        }
    }
    free(fill);
}

static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

int64_t or_symmetrize(int64_t n, const int64_t *off, const int32_t *adj,
                      int64_t *off_out, int32_t *adj_out) {
    int64_t m = off[n];
    int64_t *toff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t *tadj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    or_transpose(n, m, off, adj, toff, tadj, NULL);
    int32_t *buf = NULL;
    int64_t bufcap = 0;
    int64_t total = 0;
    for (int64_t v = 0; v < n; v++) {
        int64_t d1 = off[v + 1] - off[v], d2 = toff[v + 1] - toff[v];
        if (d1 + d2 > bufcap) {
            bufcap = (d1 + d2) * 2;
            buf = (int32_t *)realloc(buf, sizeof(int32_t) * (size_t)bufcap);
        }
        memcpy(buf, adj + off[v], sizeof(int32_t) * (size_t)d1);
        memcpy(buf + d1, tadj + toff[v], sizeof(int32_t) * (size_t)d2);
        qsort(buf, (size_t)(d1 + d2), sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t i = 0; i < d1 + d2; i++) {
            if (buf[i] == v) continue; /* no self loops */
            if (u > 0 && buf[u - 1] == buf[i]) continue;
            buf[u++] = buf[i];
        }
        if (off_out) off_out[v] = total;
        if (adj_out) memcpy(adj_out + total, buf, sizeof(int32_t) * (size_t)u);
        total += u;
    }
    if (off_out) off_out[n] = total;
    free(buf); free(toff); free(tadj);
    return total;
}

int32_t or_weight_i32(uint64_t seed, int32_t u, int32_t v) {
    uint32_t lo = (uint32_t)(u < v ? u : v), hi = (uint32_t)(u < v ? v : u);
    uint64_t h = or_mix64(or_mix64(seed ^ 0xC2B2AE3D27D4EB4Full) ^ (((uint64_t)lo << 32) | hi));
    return (int32_t)(1 + (h >> 33) % 255u);
}

float or_value_f32(uint64_t seed, int32_t u, int32_t v) {
    uint64_t h = or_mix64(or_mix64(seed ^ 0x165667B19E3779F9ull) ^
                          (((uint64_t)(uint32_t)u << 32) | (uint32_t)v));
    /* (k+1) * 2^-24 with k in [0, 2^24): exactly representable, in (0,1] */
    return (float)((h >> 40) + 1) * (1.0f / 16777216.0f);
}

float or_unit_f32(uint64_t seed, uint64_t i) {
    uint64_t h = or_mix64(or_mix64(seed ^ 0x27D4EB2F165667C5ull) + i);
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

float or_range_f32(uint64_t seed, uint64_t i, float lo, float hi) {
    float u = or_unit_f32(seed, i);
    float span = hi - lo;
    float t = span * u;   /* rounded product (no contraction: -ffp-contract=off) */
    return lo + t;
}

/* BP edge potential W_e[k/2][k%2] of the edge s->d in [lo,hi) */
float or_bp_pot(uint64_t seed, int32_t s, int32_t d, int k, float lo, float hi) {
    uint64_t i = or_mix64(((uint64_t)(uint32_t)s << 32) | (uint32_t)d) + (uint64_t)k;
    return or_range_f32(seed ^ 0xB5297A4D3F84D5B5ull, i, lo, hi);
}

void or_fill_bp_pot(uint64_t seed, int64_t n, const int64_t *off, const int32_t *adj,
                    float lo, float hi, float *pot) {
    for (int64_t s = 0; s < n; s++)
        for (int64_t e = off[s]; e < off[s + 1]; e++)
            for (int k = 0; k < 4; k++) pot[4 * e + k] = or_bp_pot(seed, (int32_t)s, adj[e], k, lo, hi);
}

void or_fill_weights_i32(uint64_t seed, int64_t n, const int64_t *off,
                         const int32_t *adj, int32_t *w) {
    for (int64_t s = 0; s < n; s++)
        for (int64_t e = off[s]; e < off[s + 1]; e++)
            w[e] = or_weight_i32(seed, (int32_t)s, adj[e]);
}

void or_fill_values_f32(uint64_t seed, int64_t n, const int64_t *off,
                        const int32_t *adj, float *w) {
    for (int64_t s = 0; s < n; s++)
        for (int64_t e = off[s]; e < off[s + 1]; e++)
            w[e] = or_value_f32(seed, (int32_t)s, adj[e]);
}

void or_fill_vertex_f32(uint64_t seed, int64_t n, float lo, float hi, float *x) {
    for (int64_t i = 0; i < n; i++) x[i] = or_range_f32(seed, (uint64_t)i, lo, hi);
}

/* ------------------------------------------------------------------ */
/* PBBS text IO                                                         */
/* ------------------------------------------------------------------ */

int or_write_adj(const char *path, int64_t n, int64_t m, const int64_t *off,
                 const int32_t *adj) {
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "AdjacencyGraph\n%lld\n%lld\n", (long long)n, (long long)m);
    for (int64_t i = 0; i < n; i++) fprintf(f, "%lld\n", (long long)off[i]);
    for (int64_t e = 0; e < m; e++) fprintf(f, "%d\n", adj[e]);
    fclose(f);
    return 0;
}

int or_write_wadj(const char *path, int64_t n, int64_t m, const int64_t *off,
                  const int32_t *adj, const int32_t *w) {
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "WeightedAdjacencyGraph\n%lld\n%lld\n", (long long)n, (long long)m);
    for (int64_t i = 0; i < n; i++) fprintf(f, "%lld\n", (long long)off[i]);
    for (int64_t e = 0; e < m; e++) fprintf(f, "%d\n", adj[e]);
    for (int64_t e = 0; e < m; e++) fprintf(f, "%d\n", w[e]);
    fclose(f);
    return 0;
}

/* The binary graph of readGraphFromBinary (IO-numa.h:543-635): <base>.config
 * holds n as text, <base>.adj the m neighbour ids as 32-bit words, <base>.idx the
 * n out-list offsets (no closing m), idx_bytes = 4 (IO-numa.h:579) or 8 (the
 * 8-byte width readWghGraphFromBinary checks for, IO-numa.h:672). */
static FILE *open_ext(const char *base, const char *ext, const char *mode) {
    char path[4096];
    if (snprintf(path, sizeof(path), "%s%s", base, ext) >= (int)sizeof(path)) return NULL;
    return fopen(path, mode);
}

int or_write_bin(const char *base, int64_t n, int64_t m, const int64_t *off,
                 const int32_t *adj, int idx_bytes) {
    FILE *f = open_ext(base, ".config", "w");
    if (!f) return -1;
    fprintf(f, "%lld\n", (long long)n);
    fclose(f);
    f = open_ext(base, ".adj", "wb");
    if (!f) return -1;
    for (int64_t e = 0; e < m; e++) {
        const uint32_t x = (uint32_t)adj[e];
        fwrite(&x, 4, 1, f);
    }
    fclose(f);
    f = open_ext(base, ".idx", "wb");
    if (!f) return -1;
    for (int64_t i = 0; i < n; i++) {
        if (idx_bytes == 8) {
            const int64_t x = off[i];
            fwrite(&x, 8, 1, f);
        } else {
            const int32_t x = (int32_t)off[i];
            fwrite(&x, 4, 1, f);
        }
    }
    fclose(f);
    return 0;
}

static int64_t file_size(FILE *f) {
    if (fseek(f, 0, SEEK_END) != 0) return -1;
    const long s = ftell(f);
    fseek(f, 0, SEEK_SET);
    return (int64_t)s;
}

/* Returns 0, or -1 (missing file), -2 (bad config), -3 (idx size neither 4n nor
 * 8n bytes: "File size wrong", where the reference aborts).  Out-degree of v is
 * off[v+1] - off[v] with off[n] = m = size(.adj) / 4 (IO-numa.h:584-590). */
int or_read_bin(const char *base, int64_t *n_out, int64_t *m_out, int64_t *off,
                int32_t *adj) {
    FILE *f = open_ext(base, ".config", "r");
    if (!f) return -1;
    long long n = 0;
    const int ok = fscanf(f, "%lld", &n) == 1 && n > 0;
    fclose(f);
    if (!ok) return -2;
    FILE *fa = open_ext(base, ".adj", "rb");
    FILE *fi = open_ext(base, ".idx", "rb");
    if (!fa || !fi) {
        if (fa) fclose(fa);
        if (fi) fclose(fi);
        return -1;
    }
    const int64_t m = file_size(fa) / 4, isz = file_size(fi);
    const int wide = isz == 8 * n;
    if (!wide && isz != 4 * n) {
        fclose(fa);
        fclose(fi);
        return -3;
    }
    *n_out = n;
    *m_out = m;
    int rc = 0;
    if (off) {
        for (int64_t i = 0; i < n && rc == 0; i++) {
            if (wide) {
                int64_t x;
                if (fread(&x, 8, 1, fi) != 1) rc = -1;
                off[i] = x;
            } else {
                int32_t x;
                if (fread(&x, 4, 1, fi) != 1) rc = -1;
                off[i] = x;
            }
        }
        off[n] = m;
    }
    if (adj && rc == 0) {
        for (int64_t e = 0; e < m && rc == 0; e++) {
            uint32_t x;
            if (fread(&x, 4, 1, fa) != 1) rc = -1;
            adj[e] = (int32_t)x;
        }
    }
    fclose(fa);
    fclose(fi);
    return rc;
}

/* Whitespace-separated tokens as in stringToWords (IO-numa.h:62-109). */
static int next_tok(FILE *f, char *buf, int cap) {
    int c, len = 0;
    do { c = fgetc(f); } while (c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0);
    if (c == EOF) return 0;
    while (c != EOF && !(c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0)) {
        if (len < cap - 1) buf[len++] = (char)c;
        c = fgetc(f);
    }
    buf[len] = 0;
    return 1;
}

int or_read_adj(const char *path, int skip_self_loops, int64_t *n_out,
                int64_t *m_out, int64_t *off, int32_t *adj, int32_t *w) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char tok[64];
    if (!next_tok(f, tok, 64)) { fclose(f); return -2; }
    int weighted;
    if (strcmp(tok, "AdjacencyGraph") == 0) weighted = 0;
    else if (strcmp(tok, "WeightedAdjacencyGraph") == 0) weighted = 1;
    else { fclose(f); return -2; }
    if (!next_tok(f, tok, 64)) { fclose(f); return -2; }
    int64_t n = atoll(tok);
    if (!next_tok(f, tok, 64)) { fclose(f); return -2; }
    int64_t m = atoll(tok);
    int64_t *o = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t *a = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int32_t *ww = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int bad = 0;
    for (int64_t i = 0; i < n && !bad; i++) { if (!next_tok(f, tok, 64)) bad = 1; else o[i] = atoll(tok); }
    for (int64_t e = 0; e < m && !bad; e++) { if (!next_tok(f, tok, 64)) bad = 1; else a[e] = (int32_t)atol(tok); }
    for (int64_t e = 0; weighted && e < m && !bad; e++) { if (!next_tok(f, tok, 64)) bad = 1; else ww[e] = (int32_t)atol(tok); }
    if (!bad && next_tok(f, tok, 64)) bad = 1; /* length check, IO-numa.h:122-126 */
    fclose(f);
    if (bad) { free(o); free(a); free(ww); return -3; }
    o[n] = m;
    int64_t mm = 0;
    for (int64_t v = 0; v < n; v++) {
        int64_t start = o[v], end = o[v + 1];
        if (off) off[v] = mm;
        for (int64_t e = start; e < end; e++) {
            if (skip_self_loops && a[e] == v) continue;
            if (adj) adj[mm] = a[e];
            if (w) w[mm] = weighted ? ww[e] : 1;
            mm++;
        }
    }
    if (off) off[n] = mm;
    *n_out = n;
    *m_out = mm;
    free(o); free(a); free(ww);
    return weighted;
}

/* ------------------------------------------------------------------ */
/* PageRankDelta (numa-PageRankDelta.C): p = 0, Delta = 1/n and every     */
/* vertex active at the start (:280-284); each round the active sources */
/* push Delta[s]/outdeg(s) into nghSum (PR_F :63-85, edgeMap            */
/* DENSE_FORWARD at threshold 0, :223); every vertex is then filtered   */
/* (:226-230): round 1 PR_Vertex_F_FirstRound (:87-102)                  */
/*   Delta = d (p + nghSum) + (1-d)/n - p;  p += Delta;  Delta -= 1/n    */
/* later rounds PR_Vertex_F (:104-117)                                   */
/*   Delta = d nghSum;  p += Delta                                       */
/* and stays active iff |Delta| > epsilon2 p; nghSum is reset (:119-127). */
/* ------------------------------------------------------------------ */

int or_pagerank_delta(int64_t n, const int64_t *off, const int32_t *adj, int iters,
                      double damping, double epsilon2, double *p, int64_t *active) {
    double *delta = (double *)malloc(sizeof(double) * (size_t)n);
    double *ngh = (double *)calloc((size_t)n, sizeof(double));
    uint8_t *F = (uint8_t *)malloc((size_t)n);
    if (!delta || !ngh || !F) { free(delta); free(ngh); free(F); return -1; }
    const double one_over_n = 1.0 / (double)n;
    const double addc = (1.0 - damping) * one_over_n;
    for (int64_t i = 0; i < n; i++) { p[i] = 0.0; delta[i] = one_over_n; F[i] = 1; }
    for (int t = 0; t < iters; t++) {
        int64_t cnt = 0;
        for (int64_t s = 0; s < n; s++) {
            if (!F[s]) continue;
            cnt++;
            const int64_t deg = off[s + 1] - off[s];
            for (int64_t e = off[s]; e < off[s + 1]; e++) ngh[adj[e]] += delta[s] / (double)deg;
        }
        if (active) active[t] = cnt;
        for (int64_t i = 0; i < n; i++) {
            if (t == 0) {
                delta[i] = damping * (p[i] + ngh[i]) + addc - p[i];
                p[i] += delta[i];
                delta[i] -= one_over_n;
            } else {
                delta[i] = ngh[i] * damping;
                p[i] += delta[i];
            }
            F[i] = fabs(delta[i]) > epsilon2 * p[i];
        }
        for (int64_t i = 0; i < n; i++) ngh[i] = 0.0;
    }
    free(delta); free(ngh); free(F);
    return 0;
}

/* ------------------------------------------------------------------ */
/* PageRank  (numa-PageRank.C:64-148 functors, :273-335 loop,           */
/*            :177-230 edgeMapDenseForwardOTHER)                        */
/* ------------------------------------------------------------------ */

int or_pagerank_push(int64_t n, const int64_t *off, const int32_t *adj,
                     int iters, double damping, double *rank_out,
                     int64_t *edges_per_iter) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)malloc(sizeof(double) * (size_t)n);
    uint8_t *F = (uint8_t *)malloc((size_t)n);
    uint8_t *N = (uint8_t *)malloc((size_t)n);
    double one_over_n = 1 / (double)n;
    double added = (1 - damping) * (1 / (double)n); /* PR_Vertex_F ctor :131 */
    for (int64_t i = 0; i < n; i++) { cur[i] = one_over_n; nxt[i] = 0; F[i] = 1; } /* :427-429 */
    for (int t = 0; t < iters; t++) {
        memset(N, 0, (size_t)n); /* clearLocalFrontier */
        int64_t et = 0;
        for (int64_t s = 0; s < n; s++) {
            if (!F[s]) continue;
            int64_t d = off[s + 1] - off[s];
            double val = cur[s];                       /* getCurrVal :214 */
            et += d;
            for (int64_t e = off[s]; e < off[s + 1]; e++) {
                nxt[adj[e]] += val / (double)d;         /* updateValVer :83-86 */
                N[adj[e]] = 1;                          /* setBit :224 */
            }
        }
        if (edges_per_iter) edges_per_iter[t] = et;
        for (int64_t i = 0; i < n; i++) if (F[i]) nxt[i] = damping * nxt[i] + added; /* :316 */
        for (int64_t i = 0; i < n; i++) if (F[i]) cur[i] = 0.0;                     /* :323 */
        double *tp = cur; cur = nxt; nxt = tp;                                       /* :326 */
        uint8_t *tf = F; F = N; N = tf;                                              /* :329 */
    }
    memcpy(rank_out, cur, sizeof(double) * (size_t)n);
    free(cur); free(nxt); free(F); free(N);
    return 0;
}

/* Same iteration as or_pagerank_push, computed by pulling over the CSC with
 * OpenMP threads.  Used as the multithreaded CPU baseline. */
int or_pagerank_pull_omp(int64_t n, const int64_t *off, const int64_t *toff,
                         const int32_t *tadj, int iters, double damping,
                         int threads, double *rank_out) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)malloc(sizeof(double) * (size_t)n);
    double *contrib = (double *)malloc(sizeof(double) * (size_t)n);
    uint8_t *F = (uint8_t *)malloc((size_t)n);
    uint8_t *N = (uint8_t *)malloc((size_t)n);
    double added = (1 - damping) * (1 / (double)n);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) { cur[i] = 1 / (double)n; nxt[i] = 0; F[i] = 1; }
    for (int t = 0; t < iters; t++) {
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < n; s++) {
            int64_t d = off[s + 1] - off[s];
            contrib[s] = (F[s] && d > 0) ? cur[s] / (double)d : 0.0;
        }
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t v = 0; v < n; v++) {
            double sum = 0.0;
            uint8_t any = 0;
            for (int64_t e = toff[v]; e < toff[v + 1]; e++) {
                int32_t s = tadj[e];
                if (F[s]) { sum += contrib[s]; any = 1; }
            }
            N[v] = any;
            if (F[v]) { nxt[v] = damping * (nxt[v] + sum) + added; cur[v] = 0.0; }
        }
        double *tp = cur; cur = nxt; nxt = tp;
        uint8_t *tf = F; F = N; N = tf;
    }
    memcpy(rank_out, cur, sizeof(double) * (size_t)n);
    free(cur); free(nxt); free(contrib); free(F); free(N);
    return 0;
}

/* ------------------------------------------------------------------ */
/* BFS  (numa-BFS.C:59-79 BFS_F, :113-159 edgeMapDenseNoRep, :258-296)  */
/* threshold 0 => every round is the dense pull (numa-BFS.C:268).       */
/* ------------------------------------------------------------------ */

int or_bfs_pull(int64_t n, const int64_t *toff, const int32_t *tadj,
                int32_t src, int32_t *parent, int32_t *level, int32_t *rounds) {
    if (src < 0 || src >= n) return -1;
    uint8_t *F = (uint8_t *)calloc((size_t)n, 1);
    uint8_t *N = (uint8_t *)calloc((size_t)n, 1);
    for (int64_t i = 0; i < n; i++) { parent[i] = -1; if (level) level[i] = -1; }
    parent[src] = src;                       /* :347-350 */
    if (level) level[src] = 0;
    F[src] = 1;
    int64_t fsize = 1;
    int r = 0;
    while (fsize > 0 || r == 0) {
        r++;
        memset(N, 0, (size_t)n);
        int64_t nsz = 0;
        for (int64_t i = 0; i < n; i++) {
            if (parent[i] != -1) continue;             /* cond */
            for (int64_t e = toff[i]; e < toff[i + 1]; e++) {
                int32_t s = tadj[e];
                if (F[s]) {                              /* getBit :150 */
                    parent[i] = s;                       /* CAS :70-72 */
                    if (level) level[i] = r;
                    N[i] = 1; nsz++;
                    break;                               /* !cond => break :153 */
                }
            }
        }
        uint8_t *t = F; F = N; N = t;
        fsize = nsz;
    }
    if (rounds) *rounds = r;
    free(F); free(N);
    return 0;
}

int or_bfs_levels(int64_t n, const int64_t *off, const int32_t *adj,
                  int32_t src, int32_t *level) {
    if (src < 0 || src >= n) return -1;
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) level[i] = -1;
    int64_t h = 0, t = 0;
    q[t++] = src; level[src] = 0;
    while (h < t) {
        int32_t v = q[h++];
        for (int64_t e = off[v]; e < off[v + 1]; e++) {
            int32_t u = adj[e];
            if (level[u] < 0) { level[u] = level[v] + 1; q[t++] = u; }
        }
    }
    free(q);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Connected components (numa-Components.C:55-119 functors, :121-175   */
/* edgeMapCustom switch at m/8, :233-281 loop; polymer.h:1097-1168      */
/* edgeMapDenseReduce, :1928-2003 edgeMapSparseV3).                     */
/* ------------------------------------------------------------------ */

int or_cc_labelprop(int64_t n, const int64_t *off, const int32_t *adj,
                    int32_t *ID, int32_t *rounds) {
    int64_t m = off[n];
    int32_t *prev = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    uint8_t *F = (uint8_t *)malloc((size_t)(n ? n : 1));
    uint8_t *N = (uint8_t *)malloc((size_t)(n ? n : 1));
    int32_t *list = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    int32_t *nlist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int64_t lsize = 0;
    int dense = 1;
    for (int64_t i = 0; i < n; i++) { ID[i] = (int32_t)i; F[i] = 1; }  /* :390, :320-325 */
    int64_t fsize = n, fedges = m;
    int64_t thr = m / 8;                                                 /* :226 */
    int r = 0;
    while (fsize > 0 || r == 0) {
        r++;
        /* vertexMap(prev = ID) over the current frontier */
        if (dense) { for (int64_t i = 0; i < n; i++) if (F[i]) prev[i] = ID[i]; }
        else { for (int64_t k = 0; k < lsize; k++) prev[list[k]] = ID[list[k]]; }
        if (fsize + fedges >= thr) {                                     /* P9: >= */
            if (!dense) { /* toDense */
                memset(F, 0, (size_t)n);
                for (int64_t k = 0; k < lsize; k++) F[list[k]] = 1;
                dense = 1;
            }
            memset(N, 0, (size_t)n);
            for (int64_t i = 0; i < n; i++) {
                int32_t data0 = ID[i], data1 = prev[i];                  /* initFunc */
                int64_t d = off[i + 1] - off[i];
                for (int64_t e = off[i]; e < off[i + 1]; e++) {
                    int32_t s = adj[e];
                    if (!F[s]) continue;
                    int32_t orig = data0;                                /* reduceFunc */
                    if (ID[s] < orig) {
                        data0 = ID[s] < orig ? ID[s] : orig;
                        if (orig == data1) N[i] = 1;
                    }
                }
                if (d > 0 && data0 < ID[i]) ID[i] = data0;               /* combineFunc writeMin */
            }
            fsize = 0; fedges = 0;                                       /* vertexCounter */
            for (int64_t i = 0; i < n; i++) if (N[i]) { fsize++; fedges += off[i + 1] - off[i]; }
            uint8_t *t = F; F = N; N = t;
        } else {
            if (dense) { /* toSparse = packIndex (ascending) */
                lsize = 0;
                for (int64_t i = 0; i < n; i++) if (F[i]) list[lsize++] = (int32_t)i;
                dense = 0;
            }
            int64_t nm = 0, ne = 0;
            for (int64_t k = 0; k < lsize; k++) {
                int32_t s = list[k];
                for (int64_t e = off[s]; e < off[s + 1]; e++) {
                    int32_t d = adj[e];
                    int32_t orig = ID[d];                                /* updateAtomic :71-75 */
                    int won = 0;
                    if (ID[s] < ID[d]) { ID[d] = ID[s]; won = 1; }
                    if (won && orig == prev[d]) { nlist[nm++] = d; ne += off[d + 1] - off[d]; }
                }
            }
            /* next list may hold duplicates in the reference as well only
             * if a vertex is appended twice; orig==prev admits one append. */
            if (nm > n) nm = n;
            memcpy(list, nlist, sizeof(int32_t) * (size_t)nm);
            lsize = nm; fsize = nm; fedges = ne;
        }
    }
    if (rounds) *rounds = r;
    free(prev); free(F); free(N); free(list); free(nlist);
    return 0;
}

static int32_t uf_find(int32_t *p, int32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

int or_cc_unionfind(int64_t n, const int64_t *off, const int32_t *adj, int32_t *label) {
    int32_t *p = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) p[i] = (int32_t)i;
    for (int64_t s = 0; s < n; s++)
        for (int64_t e = off[s]; e < off[s + 1]; e++) {
            int32_t a = uf_find(p, (int32_t)s), b = uf_find(p, adj[e]);
            if (a != b) { if (a < b) p[b] = a; else p[a] = b; }  /* root = min id */
        }
    for (int64_t i = 0; i < n; i++) label[i] = uf_find(p, (int32_t)i);
    free(p);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Bellman-Ford (numa-BellmanFord.C:61-95 functors, :166-211 loop;      */
/* polymer-wgh.h:888-949 dense push, :1274-1338 sparse, :1359-1407      */
/* switch at m/20).                                                     */
/* ------------------------------------------------------------------ */

static int sssp_bf_impl(int64_t n, const int64_t *off, const int32_t *adj,
                        const int32_t *w, int32_t src, int32_t *dist,
                        int32_t *rounds, int reset_current) {
    if (src < 0 || src >= n) return -1;
    int64_t m = off[n];
    int32_t *vis = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    uint8_t *F = (uint8_t *)calloc((size_t)n, 1);
    uint8_t *N = (uint8_t *)calloc((size_t)n, 1);
    int32_t *list = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t *nlist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int64_t lsize = 0;
    int dense = 1;
    for (int64_t i = 0; i < n; i++) dist[i] = OR_INF_DIST;             /* :283 */
    dist[src] = 0;
    F[src] = 1;
    int64_t fsize = 1, fedges = off[src + 1] - off[src];
    int64_t thr = m / 20;                                                /* :182 */
    int r = 0;
    while (fsize > 0 || r == 0) {
        r++;
        if (r > 4 * n + 8) break; /* guard against negative cycles (none here) */
        if (fsize + fedges >= thr) {
            if (!dense) {
                memset(F, 0, (size_t)n);
                for (int64_t k = 0; k < lsize; k++) F[list[k]] = 1;
                dense = 1;
            }
            memset(N, 0, (size_t)n);
            for (int64_t s = 0; s < n; s++) {
                if (!F[s]) continue;
                for (int64_t e = off[s]; e < off[s + 1]; e++) {
                    int32_t d = adj[e];
                    int32_t nd = dist[s] + w[e];
                    if (nd < dist[d]) {                                   /* writeMin */
                        dist[d] = nd;
                        if (vis[d] == 0) { vis[d] = 1; N[d] = 1; }        /* CAS(Visited) */
                    }
                }
            }
            /* reset Visited: reference = current frontier (:185, P6);
             * Ligra = next frontier. */
            if (reset_current) { for (int64_t i = 0; i < n; i++) if (F[i]) vis[i] = 0; }
            else { for (int64_t i = 0; i < n; i++) if (N[i]) vis[i] = 0; }
            fsize = 0; fedges = 0;
            for (int64_t i = 0; i < n; i++) if (N[i]) { fsize++; fedges += off[i + 1] - off[i]; }
            uint8_t *t = F; F = N; N = t;
        } else {
            if (dense) {
                lsize = 0;
                for (int64_t i = 0; i < n; i++) if (F[i]) list[lsize++] = (int32_t)i;
                dense = 0;
            }
            int64_t nm = 0, ne = 0;
            for (int64_t k = 0; k < lsize; k++) {
                int32_t s = list[k];
                for (int64_t e = off[s]; e < off[s + 1]; e++) {
                    int32_t d = adj[e];
                    int32_t nd = dist[s] + w[e];
                    if (nd < dist[d]) {
                        dist[d] = nd;
                        if (vis[d] == 0) { vis[d] = 1; nlist[nm++] = d; ne += off[d + 1] - off[d]; }
                    }
                }
            }
            if (reset_current) { for (int64_t k = 0; k < lsize; k++) vis[list[k]] = 0; }
            else { for (int64_t k = 0; k < nm; k++) vis[nlist[k]] = 0; }
            memcpy(list, nlist, sizeof(int32_t) * (size_t)nm);
            lsize = nm; fsize = nm; fedges = ne;
        }
    }
    if (rounds) *rounds = r;
    free(vis); free(F); free(N); free(list); free(nlist);
    return 0;
}

int or_sssp_bf_reference(int64_t n, const int64_t *off, const int32_t *adj,
                         const int32_t *w, int32_t src, int32_t *dist,
                         int32_t *rounds) {
    return sssp_bf_impl(n, off, adj, w, src, dist, rounds, 1);
}

int or_sssp_bf(int64_t n, const int64_t *off, const int32_t *adj,
               const int32_t *w, int32_t src, int32_t *dist, int32_t *rounds) {
    return sssp_bf_impl(n, off, adj, w, src, dist, rounds, 0);
}

/* binary heap Dijkstra */
int or_dijkstra(int64_t n, const int64_t *off, const int32_t *adj,
                const int32_t *w, int32_t src, int32_t *dist) {
    if (src < 0 || src >= n) return -1;
    int64_t m = off[n];
    int64_t cap = m + 2;
    int64_t *hk = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap); /* (dist<<32)|v */
    int64_t hs = 0;
    for (int64_t i = 0; i < n; i++) dist[i] = OR_INF_DIST;
    dist[src] = 0;
    hk[hs++] = ((int64_t)0 << 32) | src;
    while (hs > 0) {
        int64_t top = hk[0];
        hk[0] = hk[--hs];
        for (int64_t i = 0;;) { /* sift down */
            int64_t l = 2 * i + 1, r = l + 1, s = i;
            if (l < hs && hk[l] < hk[s]) s = l;
            if (r < hs && hk[r] < hk[s]) s = r;
            if (s == i) break;
            int64_t t = hk[i]; hk[i] = hk[s]; hk[s] = t; i = s;
        }
        int32_t v = (int32_t)(top & 0xFFFFFFFF);
        int32_t dv = (int32_t)(top >> 32);
        if (dv != dist[v]) continue;
        for (int64_t e = off[v]; e < off[v + 1]; e++) {
            int32_t u = adj[e];
            int32_t nd = dv + w[e];
            if (nd < dist[u]) {
                dist[u] = nd;
                if (hs >= cap) { cap *= 2; hk = (int64_t *)realloc(hk, sizeof(int64_t) * (size_t)cap); }
                int64_t i = hs++;
                hk[i] = ((int64_t)nd << 32) | u;
                while (i > 0) { int64_t p = (i - 1) / 2; if (hk[p] <= hk[i]) break; int64_t t = hk[p]; hk[p] = hk[i]; hk[i] = t; i = p; }
            }
        }
    }
    free(hk);
    return 0;
}

/* ------------------------------------------------------------------ */
/* SpMV (numa-SPMV.C:56-108 functors, :168-193 loop): the frontier All  */
/* is never switched, so every iteration is y = W^T x over all vertices.*/
/* ------------------------------------------------------------------ */

int or_spmv(int64_t n, const int64_t *off, const int32_t *adj, const double *w,
            const double *x0, int iters, double *y_out) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)calloc((size_t)n, sizeof(double));
    memcpy(cur, x0, sizeof(double) * (size_t)n);
    for (int t = 0; t < iters; t++) {
        for (int64_t s = 0; s < n; s++)
            for (int64_t e = off[s]; e < off[s + 1]; e++)
                nxt[adj[e]] += cur[s] * w[e];                 /* SPMV_F::updateAtomic */
        for (int64_t i = 0; i < n; i++) cur[i] = 0.0;          /* SPMV_Vertex_Reset */
        double *tp = cur; cur = nxt; nxt = tp;
    }
    memcpy(y_out, cur, sizeof(double) * (size_t)n);
    free(cur); free(nxt);
    return 0;
}

/* Adjoint of or_spmv (numa-SPMV.C:56-97 iterated T times, :168-193) for the
 * loss L = sum_i g_i y_T[i], y_{t+1}[d] = sum_{e=(s->d)} y_t[s] w[e]:
 *   mu_T = g,   mu_t[s] = sum_{e=(s->d)} w[e] mu_{t+1}[d]   (the transposed map),
 *   dL/dx0 = mu_0,   dL/dw[e] = sum_{t<T} y_t[s] mu_{t+1}[d]   (e = s->d).
 * fp64 throughout; the forward iterates y_0..y_{T-1} are kept.  Not part of the
 * reference (which has no derivatives): the chain rule applied to its loop. */
int or_spmv_adjoint(int64_t n, const int64_t *off, const int32_t *adj, const double *w,
                    const double *x0, const double *g, int iters, double *grad_x0,
                    double *grad_w) {
    const int64_t m = off[n];
    double *ys = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(iters + 1));
    double *mu = (double *)malloc(sizeof(double) * (size_t)n);
    double *mn = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(ys, x0, sizeof(double) * (size_t)n);
    for (int t = 0; t < iters; t++) {
        const double *y = ys + (size_t)t * n;
        double *yn = ys + (size_t)(t + 1) * n;
        for (int64_t i = 0; i < n; i++) yn[i] = 0.0;
        for (int64_t s = 0; s < n; s++)
            for (int64_t e = off[s]; e < off[s + 1]; e++) yn[adj[e]] += y[s] * w[e];
    }
    for (int64_t e = 0; e < m; e++) grad_w[e] = 0.0;
    memcpy(mu, g, sizeof(double) * (size_t)n);
    for (int t = iters - 1; t >= 0; t--) {
        const double *y = ys + (size_t)t * n;
        for (int64_t s = 0; s < n; s++) {
            double sum = 0.0;
            for (int64_t e = off[s]; e < off[s + 1]; e++) {
                sum += w[e] * mu[adj[e]];
                grad_w[e] += y[s] * mu[adj[e]];
            }
            mn[s] = sum;
        }
        double *tp = mu; mu = mn; mn = tp;
    }
    memcpy(grad_x0, mu, sizeof(double) * (size_t)n);
    free(ys); free(mu); free(mn);
    return 0;
}

/* or_spmv computed by pulling over the CSC with OpenMP threads: teid[p] is
 * the CSR edge of in-edge p.  The in-lists of or_transpose are in ascending
 * source order, which is the order in which or_spmv's push adds into each
 * destination, so every sum is formed in the same order and the result is
 * bit-identical to or_spmv (numa-SPMV.C:56-97, :168-193). */
int or_spmv_pull_omp(int64_t n, const int64_t *toff, const int32_t *tadj, const int64_t *teid,
                     const double *w, const double *x0, int iters, int threads,
                     double *y_out) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)malloc(sizeof(double) * (size_t)n);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    memcpy(cur, x0, sizeof(double) * (size_t)n);
    for (int t = 0; t < iters; t++) {
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t d = 0; d < n; d++) {
            double sum = 0.0;
            for (int64_t p = toff[d]; p < toff[d + 1]; p++) sum += cur[tadj[p]] * w[teid[p]];
            nxt[d] = sum;
        }
        double *tp = cur; cur = nxt; nxt = tp;
    }
    memcpy(y_out, cur, sizeof(double) * (size_t)n);
    free(cur); free(nxt);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Belief propagation (numa-BP.C:89-140 functors, :179-230 push,        */
/* :280-316 loop).  epot[e*4 + k*2 + l] = W_e[k][l] in CSR edge order.  */
/* The second product buffer starts at zero (mmap'd, numa-BP.C:557).    */
/* ------------------------------------------------------------------ */

static int bp_impl(int64_t n, const int64_t *off, const int32_t *adj,
                   const float *vpot, const float *epot, const double *prod0, int iters,
                   double *prod_out, int state_f32);

int or_bp(int64_t n, const int64_t *off, const int32_t *adj,
          const float *vpot, const float *epot, const double *prod0, int iters,
          double *prod_out) {
    return bp_impl(n, off, adj, vpot, epot, prod0, iters, prod_out, 0);
}

/* fp64 arithmetic on a state rounded to fp32 after every iteration: the error
 * left between this and an fp32 implementation is one iteration's arithmetic
 * (the state's own rounding is raised to ~in-degree power by the next
 * iteration, which no fp32 state can avoid). */
int or_bp_state_f32(int64_t n, const int64_t *off, const int32_t *adj,
                    const float *vpot, const float *epot, const double *prod0, int iters,
                    double *prod_out) {
    return bp_impl(n, off, adj, vpot, epot, prod0, iters, prod_out, 1);
}

static int bp_impl(int64_t n, const int64_t *off, const int32_t *adj,
                   const float *vpot, const float *epot, const double *prod0, int iters,
                   double *prod_out, int state_f32) {
    double *cur = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *nxt = (double *)calloc(2 * (size_t)n, sizeof(double));
    uint8_t *F = (uint8_t *)malloc((size_t)n);
    uint8_t *N = (uint8_t *)malloc((size_t)n);
    memcpy(cur, prod0, sizeof(double) * 2 * (size_t)n);
    for (int64_t i = 0; i < n; i++) F[i] = 1;
    for (int t = 0; t < iters; t++) {
        memset(N, 0, (size_t)n);
        for (int64_t i = 0; i < n; i++) if (F[i]) { nxt[2 * i] = 1.0; nxt[2 * i + 1] = 1.0; }
        for (int64_t s = 0; s < n; s++) {
            if (!F[s]) continue;
            for (int64_t e = off[s]; e < off[s + 1]; e++) {
                int32_t d = adj[e];
                for (int k = 0; k < 2; k++) {
                    double b = 0.0;
                    for (int l = 0; l < 2; l++)
                        b += (double)vpot[2 * d + l] * (double)epot[4 * e + 2 * k + l] * cur[2 * d + l];
                    nxt[2 * d + k] *= b;
                }
                N[d] = 1;
            }
        }
        if (state_f32)
            for (int64_t i = 0; i < 2 * n; i++) nxt[i] = (double)(float)nxt[i];
        double *tp = cur; cur = nxt; nxt = tp;
        uint8_t *tf = F; F = N; N = tf;
    }
    memcpy(prod_out, cur, sizeof(double) * 2 * (size_t)n);
    free(cur); free(nxt); free(F); free(N);
    return 0;
}

/* or_bp computed by pulling over the CSC with OpenMP threads (teid[p]: the
 * CSR edge of in-edge p).  A destination multiplies the messages of its active
 * in-neighbours in ascending source order, the order in which or_bp's push
 * multiplies them, with the same expression for each message, so the result is
 * bit-identical to or_bp; a destination reached by no active source keeps its
 * product (1, 1) when active and the other buffer's content otherwise, and the
 * next frontier is the set of reached destinations (numa-BP.C:280-316). */
int or_bp_pull_omp(int64_t n, const int64_t *toff, const int32_t *tadj, const int64_t *teid,
                   const float *vpot, const float *epot, const double *prod0, int iters,
                   int threads, double *prod_out) {
    double *cur = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *nxt = (double *)calloc(2 * (size_t)n, sizeof(double));
    uint8_t *F = (uint8_t *)malloc((size_t)n);
    uint8_t *N = (uint8_t *)malloc((size_t)n);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    memcpy(cur, prod0, sizeof(double) * 2 * (size_t)n);
    for (int64_t i = 0; i < n; i++) F[i] = 1;
    for (int t = 0; t < iters; t++) {
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t d = 0; d < n; d++) {
            double p0 = F[d] ? 1.0 : nxt[2 * d], p1 = F[d] ? 1.0 : nxt[2 * d + 1];
            uint8_t any = 0;
            for (int64_t q = toff[d]; q < toff[d + 1]; q++) {
                if (!F[tadj[q]]) continue;
                const int64_t e = teid[q];
                double b0 = 0.0, b1 = 0.0;
                for (int l = 0; l < 2; l++)
                    b0 += (double)vpot[2 * d + l] * (double)epot[4 * e + l] * cur[2 * d + l];
                p0 *= b0;
                for (int l = 0; l < 2; l++)
                    b1 += (double)vpot[2 * d + l] * (double)epot[4 * e + 2 + l] * cur[2 * d + l];
                p1 *= b1;
                any = 1;
            }
            nxt[2 * d] = p0;
            nxt[2 * d + 1] = p1;
            N[d] = any;
        }
        double *tp = cur; cur = nxt; nxt = tp;
        uint8_t *tf = F; F = N; N = tf;
    }
    memcpy(prod_out, cur, sizeof(double) * 2 * (size_t)n);
    free(cur); free(nxt); free(F); free(N);
    return 0;
}

int or_bp_f32(int64_t n, const int64_t *off, const int32_t *adj,
              const float *vpot, const float *epot, const float *prod0,
              int iters, float *prod_out) {
    float *cur = (float *)malloc(sizeof(float) * 2 * (size_t)n);
    float *nxt = (float *)calloc(2 * (size_t)n, sizeof(float));
    uint8_t *F = (uint8_t *)malloc((size_t)n);
    uint8_t *N = (uint8_t *)malloc((size_t)n);
    memcpy(cur, prod0, sizeof(float) * 2 * (size_t)n);
    for (int64_t i = 0; i < n; i++) F[i] = 1;
    for (int t = 0; t < iters; t++) {
        memset(N, 0, (size_t)n);
        for (int64_t i = 0; i < n; i++) if (F[i]) { nxt[2 * i] = 1.0f; nxt[2 * i + 1] = 1.0f; }
        for (int64_t s = 0; s < n; s++) {
            if (!F[s]) continue;
            for (int64_t e = off[s]; e < off[s + 1]; e++) {
                int32_t d = adj[e];
                for (int k = 0; k < 2; k++) {
                    float b = 0.0f;
                    for (int l = 0; l < 2; l++)
                        b += vpot[2 * d + l] * epot[4 * e + 2 * k + l] * cur[2 * d + l];
                    nxt[2 * d + k] = nxt[2 * d + k] * b;
                }
                N[d] = 1;
            }
        }
        float *tp = cur; cur = nxt; nxt = tp;
        uint8_t *tf = F; F = N; N = tf;
    }
    memcpy(prod_out, cur, sizeof(float) * 2 * (size_t)n);
    free(cur); free(nxt); free(F); free(N);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Personalized PageRank (all vertices active every iteration):         */
/* r_0 = v, r_{t+1}[d] = damping * sum_{s->d} r_t[s]/outdeg(s)          */
/*                       + (1-damping) * v[d].                          */
/* ------------------------------------------------------------------ */

int or_ppr(int64_t n, const int64_t *off, const int32_t *adj, int iters,
           double damping, const double *tele, double *rank_out) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(cur, tele, sizeof(double) * (size_t)n);
    for (int t = 0; t < iters; t++) {
        for (int64_t i = 0; i < n; i++) nxt[i] = 0.0;
        for (int64_t s = 0; s < n; s++) {
            int64_t d = off[s + 1] - off[s];
            if (d == 0) continue;
            double c = cur[s] / (double)d;
            for (int64_t e = off[s]; e < off[s + 1]; e++) nxt[adj[e]] += c;
        }
        for (int64_t i = 0; i < n; i++) nxt[i] = damping * nxt[i] + (1.0 - damping) * tele[i];
        double *tp = cur; cur = nxt; nxt = tp;
    }
    memcpy(rank_out, cur, sizeof(double) * (size_t)n);
    free(cur); free(nxt);
    return 0;
}

/* Weighted PageRank x = damping * W x + (1-damping) * v with W[d][s] =
 * wn[e] for CSR edge e = (s->d); fixed-point iteration from x_0 = v until
 * ||x_{t+1} - x_t||_1 <= tol. */
int or_wpr_forward(int64_t n, const int64_t *off, const int32_t *adj,
                   const double *wn, const double *tele, double damping,
                   int max_iters, double tol, double *x_out, int *iters_used) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(cur, tele, sizeof(double) * (size_t)n);
    int t;
    for (t = 0; t < max_iters; t++) {
        for (int64_t i = 0; i < n; i++) nxt[i] = 0.0;
        for (int64_t s = 0; s < n; s++)
            for (int64_t e = off[s]; e < off[s + 1]; e++) nxt[adj[e]] += wn[e] * cur[s];
        double res = 0.0;
        for (int64_t i = 0; i < n; i++) {
            nxt[i] = damping * nxt[i] + (1.0 - damping) * tele[i];
            res += fabs(nxt[i] - cur[i]);
        }
        double *tp = cur; cur = nxt; nxt = tp;
        if (res <= tol) { t++; break; }
    }
    if (iters_used) *iters_used = t;
    memcpy(x_out, cur, sizeof(double) * (size_t)n);
    free(cur); free(nxt);
    return 0;
}

/* Adjoint of or_wpr_forward for loss g^T x:
 *   lambda = damping * W^T lambda + g   (same fixed-point iteration on the
 *   transposed graph), dL/dW_e = damping * lambda[d] * x[s],
 *   dL/dv = (1-damping) * lambda. */
int or_wpr_adjoint(int64_t n, const int64_t *off, const int32_t *adj,
                   const double *wn, const double *x, const double *g,
                   double damping, int max_iters, double tol,
                   double *lambda_out, double *grad_w, double *grad_tele,
                   int *iters_used) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)n);
    double *nxt = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(cur, g, sizeof(double) * (size_t)n);
    int t;
    for (t = 0; t < max_iters; t++) {
        double res = 0.0;
        for (int64_t s = 0; s < n; s++) {
            double acc = 0.0;
            for (int64_t e = off[s]; e < off[s + 1]; e++) acc += wn[e] * cur[adj[e]];
            nxt[s] = damping * acc + g[s];
            res += fabs(nxt[s] - cur[s]);
        }
        double *tp = cur; cur = nxt; nxt = tp;
        if (res <= tol) { t++; break; }
    }
    if (iters_used) *iters_used = t;
    if (lambda_out) memcpy(lambda_out, cur, sizeof(double) * (size_t)n);
    if (grad_w)
        for (int64_t s = 0; s < n; s++)
            for (int64_t e = off[s]; e < off[s + 1]; e++) grad_w[e] = damping * cur[adj[e]] * x[s];
    if (grad_tele)
        for (int64_t i = 0; i < n; i++) grad_tele[i] = (1.0 - damping) * cur[i];
    free(cur); free(nxt);
    return 0;
}
