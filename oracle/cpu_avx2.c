/*
 * cpu_avx2.c -- CPU baseline: restatement of par2-rs's selected x86 strategy.
 *
 * TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and the oracle tests).
 *
 * Algorithm restated (own code, same steps):
 *  - "shuffle2x" layout: in each 32-byte block of 16 words the 16 low bytes sit
 *    in the lower 128-bit lane and the 16 high bytes in the upper lane
 *    (src/galois/simd/x86/avx2.rs:571-630, prepare/finish);
 *  - per coefficient, four 16-entry nibble tables split into "same lane" and
 *    "cross lane" halves (avx2.rs:185-250, from the MUL128 nibble LUT of
 *    src/galois/core.rs:116-163);
 *  - multiply-add per 32 bytes = 4 pshufb + a lane swap + xors
 *    (avx2.rs:632-677, native shuffle2x muladd);
 *  - StreamingEncoder batching: data batches of clamp(ncpu*{2,4,8},16,128)
 *    slices, recovery slices split into groups of clamp(R/ncpu,4,8) processed in
 *    parallel, each group sweeping every slice of the batch
 *    (src/par2_rs.rs:88-126, :816-870).
 */
#include "par2_oracle.h"

#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int po_have_avx2(void) { return __builtin_cpu_supports("avx2"); }

typedef struct {
    __m256i same_lo, cross_lo, same_hi, cross_hi;
} s2x_tables;

/* T_i[v] = c * (v << 4i) for nibble position i; split into low/high bytes. */
static s2x_tables make_tables(uint16_t c) {
    uint8_t lo[4][16], hi[4][16];
    for (int i = 0; i < 4; i++)
        for (int v = 0; v < 16; v++) {
            uint16_t p = po_gf_mul((uint16_t)(v << (4 * i)), c);
            lo[i][v] = (uint8_t)p;
            hi[i][v] = (uint8_t)(p >> 8);
        }
    s2x_tables t;
    /* lower lane processes input low bytes (nibbles 0,1), upper lane the high
     * bytes (nibbles 2,3); "same" tables produce the byte that stays in the lane */
    t.same_lo = _mm256_setr_m128i(_mm_loadu_si128((const __m128i *)lo[0]), _mm_loadu_si128((const __m128i *)hi[2]));
    t.cross_lo = _mm256_setr_m128i(_mm_loadu_si128((const __m128i *)hi[0]), _mm_loadu_si128((const __m128i *)lo[2]));
    t.same_hi = _mm256_setr_m128i(_mm_loadu_si128((const __m128i *)lo[1]), _mm_loadu_si128((const __m128i *)hi[3]));
    t.cross_hi = _mm256_setr_m128i(_mm_loadu_si128((const __m128i *)hi[1]), _mm_loadu_si128((const __m128i *)lo[3]));
    return t;
}

/* interleaved u16 -> shuffle2x, in place (whole 16-word blocks only) */
static void to_s2x(uint16_t *d, size_t words) {
    const __m256i sep = _mm256_setr_epi8(0, 2, 4, 6, 8, 10, 12, 14, 1, 3, 5, 7, 9, 11, 13, 15,
                                         0, 2, 4, 6, 8, 10, 12, 14, 1, 3, 5, 7, 9, 11, 13, 15);
    for (size_t i = 0; i + 16 <= words; i += 16) {
        __m256i v = _mm256_loadu_si256((const __m256i *)(d + i));
        v = _mm256_shuffle_epi8(v, sep);
        v = _mm256_permute4x64_epi64(v, 0xD8); /* qwords 0,2,1,3 */
        _mm256_storeu_si256((__m256i *)(d + i), v);
    }
}

static void from_s2x(uint16_t *d, size_t words) {
    const __m256i join = _mm256_setr_epi8(0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15,
                                          0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15);
    for (size_t i = 0; i + 16 <= words; i += 16) {
        __m256i v = _mm256_loadu_si256((const __m256i *)(d + i));
        v = _mm256_permute4x64_epi64(v, 0xD8);
        v = _mm256_shuffle_epi8(v, join);
        _mm256_storeu_si256((__m256i *)(d + i), v);
    }
}

/* dst ^= c * src, both in shuffle2x layout; tail (<16 words) interleaved. */
static void muladd_s2x(uint16_t *dst, const uint16_t *src,
                       size_t words, uint16_t c, const s2x_tables *t) {
    if (c == 0) return;
    const __m256i m = _mm256_set1_epi8(0x0f);
    size_t i = 0;
    for (; i + 16 <= words; i += 16) {
        __m256i s = _mm256_loadu_si256((const __m256i *)(src + i));
        __m256i nl = _mm256_and_si256(s, m);
        __m256i nh = _mm256_and_si256(_mm256_srli_epi16(s, 4), m);
        __m256i same = _mm256_xor_si256(_mm256_shuffle_epi8(t->same_lo, nl), _mm256_shuffle_epi8(t->same_hi, nh));
        __m256i cross = _mm256_xor_si256(_mm256_shuffle_epi8(t->cross_lo, nl), _mm256_shuffle_epi8(t->cross_hi, nh));
        cross = _mm256_permute2x128_si256(cross, cross, 0x01);
        __m256i d = _mm256_loadu_si256((const __m256i *)(dst + i));
        _mm256_storeu_si256((__m256i *)(dst + i), _mm256_xor_si256(d, _mm256_xor_si256(same, cross)));
    }
    for (; i < words; i++) dst[i] ^= po_gf_mul(src[i], c);
}

typedef struct {
    const uint16_t *coeff; /* [r][n] */
    size_t n, r, words;
    uint16_t *rec;         /* [r][words], shuffle2x layout while encoding */
    uint16_t **batch;      /* prepared input slices of the current batch */
    const size_t *batch_idx;
    size_t batch_len, rgroup;
    s2x_tables *tabs;      /* [r][n] tables, built lazily per batch */
} enc_job;

static void flush_group(void *p, size_t g) {
    enc_job *e = p;
    size_t r0 = g * e->rgroup, r1 = r0 + e->rgroup;
    if (r1 > e->r) r1 = e->r;
    for (size_t s = 0; s < e->batch_len; s++) {
        size_t d = e->batch_idx[s];
        for (size_t i = r0; i < r1; i++)
            muladd_s2x(e->rec + i * e->words, e->batch[s], e->words, e->coeff[i * e->n + d],
                       &e->tabs[i * e->n + d]);
    }
}

typedef struct { enc_job *e; size_t base; } tab_job;
static void build_tabs(void *p, size_t k) {
    tab_job *t = p;
    enc_job *e = t->e;
    size_t d = e->batch_idx[k];
    for (size_t i = 0; i < e->r; i++) e->tabs[i * e->n + d] = make_tables(e->coeff[i * e->n + d]);
}

typedef struct { uint16_t **bufs; const uint16_t *data; const size_t *idx; size_t words; } prep_job;
static void prep_one(void *p, size_t s) {
    prep_job *j = p;
    memcpy(j->bufs[s], j->data + j->idx[s] * j->words, j->words * sizeof(uint16_t));
    to_s2x(j->bufs[s], j->words);
}

/* minimal dynamic parallel-for (duplicated from par2_oracle.c to keep this unit standalone) */
typedef struct { void (*fn)(void *, size_t); void *arg; size_t next, count; pthread_mutex_t mu; } pj;
static void *pw(void *p) {
    pj *j = p;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        size_t i = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->count) return NULL;
        j->fn(j->arg, i);
    }
}
static void pfor(int threads, size_t count, void (*fn)(void *, size_t), void *arg) {
    pj j = {fn, arg, 0, count, PTHREAD_MUTEX_INITIALIZER};
    if ((size_t)threads > count) threads = (int)count;
    if (threads <= 1) { for (size_t i = 0; i < count; i++) fn(arg, i); return; }
    pthread_t t[256];
    if (threads > 256) threads = 256;
    for (int i = 0; i < threads; i++) pthread_create(&t[i], NULL, pw, &j);
    for (int i = 0; i < threads; i++) pthread_join(t[i], NULL);
}

int po_encode_avx2(const uint16_t *data, size_t n, size_t words,
                   const uint32_t *exps, size_t r, uint16_t *rec, int threads) {
    if (!po_have_avx2()) return -1;
    po_init();
    if (threads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 0 ? (int)c : 1;
    }
    size_t ncpu = (size_t)threads;
    /* src/par2_rs.rs:88-108 and :111-126 */
    size_t batch = r >= 100 ? ncpu * 8 : r >= 50 ? ncpu * 4 : ncpu * 2;
    if (batch < 16) batch = 16;
    if (batch > 128) batch = 128;
    size_t rgroup = r / ncpu;
    if (rgroup < 1) rgroup = 1;
    if (rgroup < 4) rgroup = 4;
    if (rgroup > 8) rgroup = 8;

    uint16_t *consts = malloc(n * sizeof *consts);
    uint16_t *coeff = malloc(n * r * sizeof *coeff);
    s2x_tables *tabs = aligned_alloc(32, ((n * r * sizeof(s2x_tables)) + 31) / 32 * 32);
    po_par2_constants(n, consts);
    po_coeff_matrix(consts, n, exps, r, coeff);
    memset(rec, 0, r * words * sizeof *rec);

    uint16_t **bufs = malloc(batch * sizeof *bufs);
    for (size_t s = 0; s < batch; s++) bufs[s] = aligned_alloc(64, (words * 2 + 63) / 64 * 64);
    size_t *idx = malloc(batch * sizeof *idx);

    enc_job e = {coeff, n, r, words, rec, bufs, idx, 0, rgroup, tabs};
    for (size_t d0 = 0; d0 < n; d0 += batch) {
        size_t len = n - d0 < batch ? n - d0 : batch;
        for (size_t s = 0; s < len; s++) idx[s] = d0 + s;
        prep_job pj_ = {bufs, data, idx, words};
        pfor(threads, len, prep_one, &pj_);
        e.batch_len = len;
        tab_job tj = {&e, d0};
        pfor(threads, len, build_tabs, &tj);
        pfor(threads, (r + rgroup - 1) / rgroup, flush_group, &e);
    }
    for (size_t i = 0; i < r; i++) from_s2x(rec + i * words, words);

    for (size_t s = 0; s < batch; s++) free(bufs[s]);
    free(bufs);
    free(idx);
    free(tabs);
    free(coeff);
    free(consts);
    return 0;
}
