/*
 * par2_oracle.c -- CPU restatement of par2-rs's Reed-Solomon + slice-hash path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  See par2_oracle.h
 * for the function list; each definition cites the reference file:line it
 * restates.  Parity pins for every function live in tests/test_oracle.py.
 */
#include "par2_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define GF_SIZE 65536u
#define GF_ORDER 65535u
#define PRIMITIVE_POLY 0x1100Bu /* src/galois/core.rs:12 */

static uint16_t g_log[GF_SIZE];
static uint16_t g_exp[2 * GF_ORDER];
static int g_ready = 0;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

/* src/galois/core.rs:49-91: walk b = 2^i, doubling with xor 0x1100B on
 * overflow; exp is stored twice so exp[la+lb] needs no reduction; log[0]=0. */
static void init_tables_once(void) {
    uint32_t b = 1;
    for (uint32_t i = 0; i < GF_ORDER; i++) {
        g_exp[i] = (uint16_t)b;
        g_exp[i + GF_ORDER] = (uint16_t)b;
        g_log[b] = (uint16_t)i;
        b <<= 1;
        if (b & 0x10000u) b ^= PRIMITIVE_POLY;
    }
    g_log[0] = 0;
    g_ready = 1;
}

void po_init(void) { pthread_once(&g_once, init_tables_once); }

void po_tables(uint16_t *log_out, uint16_t *exp_out) {
    po_init();
    if (log_out) memcpy(log_out, g_log, sizeof g_log);
    if (exp_out) memcpy(exp_out, g_exp, sizeof g_exp);
}

/* src/galois/core.rs:175-189 */
uint16_t po_gf_mul(uint16_t a, uint16_t b) {
    if (!g_ready) po_init();
    if (a == 0 || b == 0) return 0;
    return g_exp[(uint32_t)g_log[a] + g_log[b]];
}

/* src/galois/core.rs:205-230 */
int po_gf_div(uint16_t a, uint16_t b, uint16_t *q) {
    if (!g_ready) po_init();
    if (b == 0) return -1;
    if (a == 0) { *q = 0; return 0; }
    uint32_t la = g_log[a], lb = g_log[b];
    *q = g_exp[la >= lb ? la - lb : la + GF_ORDER - lb];
    return 0;
}

/* src/galois/core.rs:242-258: 0^0 = 1, 0^n = 0, a^n = exp[(log a * n) mod 65535] */
uint16_t po_gf_pow(uint16_t a, uint64_t n) {
    if (!g_ready) po_init();
    if (a == 0) return n == 0 ? 1 : 0;
    if (n == 0) return 1;
    uint64_t l = ((uint64_t)g_log[a] * n) % GF_ORDER;
    return g_exp[l];
}

/* src/par2_rs.rs:64-76: n = 1, 2, ... skipping multiples of 3, 5, 17, 257 */
void po_par2_constants(size_t count, uint16_t *out) {
    uint64_t n = 1;
    size_t k = 0;
    while (k < count) {
        if (n % 3 && n % 5 && n % 17 && n % 257) out[k++] = po_gf_pow(2, n);
        n++;
    }
}

/* src/par2_rs.rs:707-719 (c_d^r by repeated multiply, starting at 1) and
 * :353-359 / :395-402 (gf_pow for arbitrary recovery exponents). */
void po_coeff_matrix(const uint16_t *constants, size_t n, const uint32_t *exps,
                     size_t r, uint16_t *coeff) {
    int contiguous = 1;
    for (size_t i = 0; i < r; i++)
        if (exps[i] != i) { contiguous = 0; break; }
    for (size_t d = 0; d < n; d++) {
        if (contiguous) {
            uint16_t c = 1;
            for (size_t i = 0; i < r; i++) {
                coeff[i * n + d] = c;
                c = po_gf_mul(c, constants[d]);
            }
        } else {
            for (size_t i = 0; i < r; i++) coeff[i * n + d] = po_gf_pow(constants[d], exps[i]);
        }
    }
}

/* src/galois/scalar.rs:9-17 */
void po_mul_slice(uint16_t k, uint16_t *data, size_t len) {
    for (size_t i = 0; i < len; i++) data[i] = po_gf_mul(data[i], k);
}

/* src/galois/scalar.rs:21-28 */
void po_muladd(uint16_t *dst, const uint16_t *src, size_t len, uint16_t k) {
    for (size_t i = 0; i < len; i++) dst[i] ^= po_gf_mul(src[i], k);
}

/* src/galois/scalar.rs:82-113 */
void po_muladd_region(uint16_t *const *dsts, size_t ndst, const uint16_t *const *srcs,
                      size_t nsrc, const uint16_t *coeff, size_t coeff_stride,
                      size_t off, size_t len) {
    for (size_t i = 0; i < ndst; i++) {
        uint16_t *dst = dsts[i] + off;
        for (size_t j = 0; j < nsrc; j++) {
            uint16_t c = coeff[i * coeff_stride + j];
            if (c == 0) continue;
            const uint16_t *src = srcs[j] + off;
            for (size_t w = 0; w < len; w++) dst[w] ^= po_gf_mul(src[w], c);
        }
    }
}

/* src/par2_rs.rs:785-902: every recovery slice starts at zero and receives
 * coeff[r][d] * data[d] for every input slice (batching changes only the
 * order of the XORs, which commute). */
void po_encode(const uint16_t *data, size_t n, size_t words,
               const uint32_t *exps, size_t r, uint16_t *rec) {
    po_init();
    uint16_t *consts = malloc(n * sizeof *consts);
    uint16_t *coeff = malloc(n * r * sizeof *coeff);
    po_par2_constants(n, consts);
    po_coeff_matrix(consts, n, exps, r, coeff);
    memset(rec, 0, r * words * sizeof *rec);
    /* log-domain inner loop: identical products to gf_mul, just faster */
    for (size_t i = 0; i < r; i++) {
        uint16_t *dst = rec + i * words;
        for (size_t d = 0; d < n; d++) {
            uint16_t c = coeff[i * n + d];
            if (c == 0) continue;
            uint32_t lc = g_log[c];
            const uint16_t *src = data + d * words;
            for (size_t w = 0; w < words; w++) {
                uint16_t s = src[w];
                if (s) dst[w] ^= g_exp[g_log[s] + lc];
            }
        }
    }
    free(consts);
    free(coeff);
}

/* src/par2_rs.rs:319-422 */
int po_transform(const uint16_t *constants, const size_t *missing, size_t m,
                 const size_t *present, size_t npresent, const uint32_t *rec_exps,
                 size_t *row_order, uint16_t *scale, uint16_t *elim, uint16_t *vand) {
    if (m == 0) return -2;
    uint16_t *A = malloc(m * m * sizeof *A);
    for (size_t row = 0; row < m; row++)
        for (size_t col = 0; col < m; col++)
            A[row * m + col] = po_gf_pow(constants[missing[col]], rec_exps[row]);
    for (size_t i = 0; i < m; i++) { row_order[i] = i; scale[i] = 1; }
    memset(elim, 0, m * m * sizeof *elim);
    uint16_t *tmp = malloc(m * sizeof *tmp);
    for (size_t k = 0; k < m; k++) {
        if (A[k * m + k] == 0) {
            size_t r;
            for (r = k + 1; r < m; r++)
                if (A[r * m + k] != 0) break;
            if (r == m) { free(A); free(tmp); return -1; }
            /* the whole row state (matrix row, order, scale, elimination row) swaps */
            memcpy(tmp, A + k * m, m * sizeof *tmp);
            memcpy(A + k * m, A + r * m, m * sizeof *tmp);
            memcpy(A + r * m, tmp, m * sizeof *tmp);
            size_t t = row_order[k]; row_order[k] = row_order[r]; row_order[r] = t;
            uint16_t s = scale[k]; scale[k] = scale[r]; scale[r] = s;
            memcpy(tmp, elim + k * m, m * sizeof *tmp);
            memcpy(elim + k * m, elim + r * m, m * sizeof *tmp);
            memcpy(elim + r * m, tmp, m * sizeof *tmp);
        }
        uint16_t pivot = A[k * m + k];
        scale[k] = pivot;
        if (pivot != 1)
            for (size_t j = k; j < m; j++)
                if (A[k * m + j]) po_gf_div(A[k * m + j], pivot, &A[k * m + j]);
        for (size_t r = 0; r < m; r++) {
            if (r == k) continue;
            uint16_t f = A[r * m + k];
            elim[r * m + k] = f;
            if (f == 0) continue;
            for (size_t j = k; j < m; j++)
                if (A[k * m + j]) A[r * m + j] ^= po_gf_mul(A[k * m + j], f);
        }
    }
    for (size_t i = 0; i < m; i++)
        for (size_t j = 0; j < npresent; j++)
            vand[i * npresent + j] = po_gf_pow(constants[present[j]], rec_exps[i]);
    free(A);
    free(tmp);
    return 0;
}

/* src/par2_rs.rs:442-675.  Faithful, including the order rule the reference
 * uses: recovery chunk i lands in accumulator row_order^-1(i) (:488-498) but is
 * combined with vand row `acc index` (:564-584); with no pivot swap both agree. */
void po_reconstruct_chunk(size_t m, size_t npresent, const size_t *row_order,
                          const uint16_t *scale, const uint16_t *elim, const uint16_t *vand,
                          const uint16_t *const *rec_chunks, const uint16_t *const *data_chunks,
                          size_t words, uint16_t *const *out) {
    uint16_t **acc = malloc(m * sizeof *acc);
    for (size_t i = 0; i < m; i++) acc[i] = calloc(words ? words : 1, sizeof(uint16_t));
    for (size_t orig = 0; orig < m; orig++) {
        size_t swapped = 0;
        while (row_order[swapped] != orig) swapped++;
        memcpy(acc[swapped], rec_chunks[orig], words * sizeof(uint16_t));
    }
    /* acc[i] ^= vand[i][j] * data_j (region muladd over all present columns) */
    for (size_t i = 0; i < m; i++)
        po_muladd_region(&acc[i], 1, data_chunks, npresent, vand + i * npresent, npresent, 0, words);
    /* replay of the elimination, :618-655 */
    for (size_t k = 0; k < m; k++) {
        uint16_t pivot = scale[k];
        if (pivot != 1)
            for (size_t s = 0; s < words; s++)
                if (acc[k][s]) po_gf_div(acc[k][s], pivot, &acc[k][s]);
        for (size_t r = 0; r < k; r++) {
            uint16_t f = elim[r * m + k];
            if (f) po_muladd(acc[r], acc[k], words, f);
        }
        for (size_t r = k + 1; r < m; r++) {
            uint16_t f = elim[r * m + k];
            if (f) po_muladd(acc[r], acc[k], words, f);
        }
    }
    for (size_t i = 0; i < m; i++) {
        memcpy(out[i], acc[i], words * sizeof(uint16_t));
        free(acc[i]);
    }
    free(acc);
}

/* Plain Gauss-Jordan on [A | I] with "first nonzero row below" pivoting. */
int po_gf_invert(const uint16_t *a, size_t m, uint16_t *inv) {
    uint16_t *A = malloc(m * m * sizeof *A);
    memcpy(A, a, m * m * sizeof *A);
    memset(inv, 0, m * m * sizeof *inv);
    for (size_t i = 0; i < m; i++) inv[i * m + i] = 1;
    for (size_t k = 0; k < m; k++) {
        size_t p = k;
        while (p < m && A[p * m + k] == 0) p++;
        if (p == m) { free(A); return -1; }
        if (p != k)
            for (size_t j = 0; j < m; j++) {
                uint16_t t = A[k * m + j]; A[k * m + j] = A[p * m + j]; A[p * m + j] = t;
                t = inv[k * m + j]; inv[k * m + j] = inv[p * m + j]; inv[p * m + j] = t;
            }
        uint16_t piv_inv;
        po_gf_div(1, A[k * m + k], &piv_inv);
        for (size_t j = 0; j < m; j++) {
            A[k * m + j] = po_gf_mul(A[k * m + j], piv_inv);
            inv[k * m + j] = po_gf_mul(inv[k * m + j], piv_inv);
        }
        for (size_t r = 0; r < m; r++) {
            if (r == k) continue;
            uint16_t f = A[r * m + k];
            if (!f) continue;
            for (size_t j = 0; j < m; j++) {
                A[r * m + j] ^= po_gf_mul(A[k * m + j], f);
                inv[r * m + j] ^= po_gf_mul(inv[k * m + j], f);
            }
        }
    }
    free(A);
    return 0;
}

/* ------------------------------------------------------------------ MD5 --
 * RFC 1321 section 3.4: four rounds of 16 steps; T[i] = floor(2^32 |sin(i+1)|). */
static uint32_t md5_T[64];
static const uint8_t md5_S[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                                  5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                                  4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                                  6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};
static pthread_once_t md5_once = PTHREAD_ONCE_INIT;
static void md5_init_T(void) {
    for (int i = 0; i < 64; i++) md5_T[i] = (uint32_t)(uint64_t)floor(fabs(sin((double)(i + 1))) * 4294967296.0);
}

static inline uint32_t rotl32(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

static void md5_block(uint32_t st[4], const uint8_t *p) {
    uint32_t M[16];
    for (int i = 0; i < 16; i++)
        M[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
               ((uint32_t)p[4 * i + 3] << 24);
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
    for (int i = 0; i < 64; i++) {
        uint32_t f;
        int g;
        if (i < 16) { f = (b & c) | (~b & d); g = i; }
        else if (i < 32) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
        else if (i < 48) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
        else { f = c ^ (b | ~d); g = (7 * i) & 15; }
        uint32_t t = d;
        d = c;
        c = b;
        b = b + rotl32(a + f + md5_T[i] + M[g], md5_S[i]);
        a = t;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
}

/* Hash `len` bytes followed by `zpad` zero bytes (the IFSC zero padding). */
static void md5_padded(const uint8_t *data, size_t len, size_t zpad, uint8_t out[16]) {
    pthread_once(&md5_once, md5_init_T);
    uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
    size_t total = len + zpad;
    uint8_t buf[128];
    size_t pos = 0;
    while (pos + 64 <= total) {
        if (pos + 64 <= len) {
            md5_block(st, data + pos);
        } else {
            memset(buf, 0, 64);
            if (pos < len) memcpy(buf, data + pos, len - pos);
            md5_block(st, buf);
        }
        pos += 64;
    }
    size_t rem = total - pos;
    memset(buf, 0, sizeof buf);
    if (pos < len) memcpy(buf, data + pos, len - pos);
    buf[rem] = 0x80;
    size_t tail = rem < 56 ? 64 : 128;
    uint64_t bits = (uint64_t)total * 8u;
    for (int i = 0; i < 8; i++) buf[tail - 8 + i] = (uint8_t)(bits >> (8 * i));
    md5_block(st, buf);
    if (tail == 128) md5_block(st, buf + 64);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(st[i] >> (8 * j));
}

void po_md5(const uint8_t *data, size_t len, uint8_t out[16]) { md5_padded(data, len, 0, out); }

/* ---------------------------------------------------------------- CRC32 --
 * CRC-32/IEEE 802.3: reflected polynomial 0xEDB88320, init and xorout ~0. */
static uint32_t crc_tab[256];
static pthread_once_t crc_once = PTHREAD_ONCE_INIT;
static void crc_init(void) {
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t c = i;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        crc_tab[i] = c;
    }
}

uint32_t po_crc32_update(uint32_t crc, const uint8_t *data, size_t len) {
    pthread_once(&crc_once, crc_init);
    for (size_t i = 0; i < len; i++) crc = crc_tab[(crc ^ data[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}

uint32_t po_crc32(const uint8_t *data, size_t len) {
    return ~po_crc32_update(0xFFFFFFFFu, data, len);
}

static uint32_t crc32_padded(const uint8_t *data, size_t len, size_t zpad) {
    static const uint8_t zeros[4096] = {0};
    uint32_t c = po_crc32_update(0xFFFFFFFFu, data, len);
    while (zpad) {
        size_t k = zpad < sizeof zeros ? zpad : sizeof zeros;
        c = po_crc32_update(c, zeros, k);
        zpad -= k;
    }
    return ~c;
}

/* src/creator.rs:389-409 (zero-pad to the slice size, then MD5 and CRC32). */
void po_hash_slices(const uint8_t *data, size_t n, size_t slice_size, size_t slice_stride,
                    const size_t *lens, uint8_t *md5_out, uint32_t *crc_out) {
    for (size_t i = 0; i < n; i++) {
        size_t len = lens ? lens[i] : slice_size;
        if (len > slice_size) len = slice_size;
        const uint8_t *p = data + i * slice_stride;
        if (md5_out) md5_padded(p, len, slice_size - len, md5_out + 16 * i);
        if (crc_out) crc_out[i] = crc32_padded(p, len, slice_size - len);
    }
}

/* --------------------------------------------- multithreaded CPU baseline -- */
typedef struct {
    void (*fn)(void *, size_t);
    void *arg;
    size_t next, count;
    pthread_mutex_t mu;
} pool_job;

static void *pool_worker(void *p) {
    pool_job *j = p;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        size_t i = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->count) break;
        j->fn(j->arg, i);
    }
    return NULL;
}

static int resolve_threads(int threads) {
    if (threads > 0) return threads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

/* Dynamic parallel-for over `count` items (rayon's work distribution). */
static void parallel_for(int threads, size_t count, void (*fn)(void *, size_t), void *arg) {
    pool_job j = {fn, arg, 0, count, PTHREAD_MUTEX_INITIALIZER};
    if (threads <= 1 || count <= 1) {
        for (size_t i = 0; i < count; i++) fn(arg, i);
        return;
    }
    if ((size_t)threads > count) threads = (int)count;
    pthread_t *t = malloc((size_t)threads * sizeof *t);
    for (int i = 0; i < threads; i++) pthread_create(&t[i], NULL, pool_worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(t[i], NULL);
    free(t);
}

typedef struct {
    const uint8_t *data;
    size_t slice_size;
    uint8_t *md5;
    uint32_t *crc;
} hash_job;

static void hash_one(void *p, size_t i) {
    hash_job *h = p;
    const uint8_t *s = h->data + i * h->slice_size;
    if (h->crc) h->crc[i] = po_crc32(s, h->slice_size);
    if (h->md5) po_md5(s, h->slice_size, h->md5 + 16 * i);
}

void po_hash_slices_mt(const uint8_t *data, size_t n, size_t slice_size,
                       uint8_t *md5_out, uint32_t *crc_out, int threads) {
    hash_job h = {data, slice_size, md5_out, crc_out};
    parallel_for(resolve_threads(threads), n, hash_one, &h);
}
