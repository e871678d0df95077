/*
 * par2_create.c -- CPU oracle (test infrastructure only): restatement of the
 * reference's PAR2 creation, producing the bytes of every output file in
 * memory from the input files' names and contents.
 *
 * Follows (paths in the par2-rs crate):
 *   block size rule          src/creator.rs:29-67   (calculate_block_size)
 *   file hashes, File ID     src/creator.rs:80-135, src/writer.rs:56-63
 *   canonical file order     src/creator.rs:280-289, src/parser.rs:556-564
 *   block / recovery counts  src/creator.rs:305-347 (f64 of the f32 percentage)
 *   IFSC of padded blocks    src/creator.rs:378-409
 *   encode, exponents 0..R-1 src/par2_rs.rs:785-902 (po_encode)
 *   main body, set ID        src/writer.rs:66-71, :137-173
 *   packet framing           src/writer.rs:83-124
 *   packet bodies            src/writer.rs:126-135, :190-264
 *   volumes and names        src/volumes.rs:47-208, src/creator.rs:453-519
 * The creator text is the literal of src/creator.rs:496.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "par2_create.h"
#include "par2_oracle.h"

static const uint8_t kMagic[8] = {'P', 'A', 'R', '2', 0, 'P', 'K', 'T'};
static const uint8_t kTypeMain[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'M', 'a', 'i', 'n', 0, 0, 0, 0};
static const uint8_t kTypeFileDesc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                          'F', 'i', 'l', 'e', 'D', 'e', 's', 'c'};
static const uint8_t kTypeIfsc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'I', 'F', 'S', 'C', 0, 0, 0, 0};
static const uint8_t kTypeRecv[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
static const uint8_t kTypeCreator[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                         'C', 'r', 'e', 'a', 't', 'o', 'r', 0};
const char *const po_par2_creator_text = "par2-rs v0.1.0 https://github.com/zephleggett/par2-rs";

/* creator.rs:29-67 */
uint64_t po_par2_block_size(uint64_t total) {
  const uint64_t kMin = 2048, kMax = 4ull * 1024 * 1024, kTarget = 2000, kMaxBlocks = 65535;
  if (total == 0) return kMin;
  uint64_t bs = total / kTarget;
  const uint64_t min_for_field = (total + kMaxBlocks - 1) / kMaxBlocks;
  if (bs % 4) bs = (bs / 4 + 1) * 4;
  uint64_t min_block = min_for_field > kMin ? min_for_field : kMin;
  if (min_block % 4) min_block = (min_block / 4 + 1) * 4;
  if (min_block > kMax) {
    bs = min_block;
  } else {
    if (bs < kMin) bs = kMin;
    if (bs > kMax) bs = kMax;
    if (bs < min_block) bs = min_block;
  }
  return bs;
}

/* writer.rs:56-63: MD5(hash_16k || length LE || name bytes) */
void po_par2_file_id(const uint8_t hash16k[16], uint64_t length, const char *name, uint8_t out[16]) {
  const size_t nl = strlen(name);
  uint8_t *buf = malloc(24 + nl);
  memcpy(buf, hash16k, 16);
  for (int i = 0; i < 8; i++) buf[16 + i] = (uint8_t)(length >> (8 * i));
  memcpy(buf + 24, name, nl);
  po_md5(buf, 24 + nl, out);
  free(buf);
}

/* parser.rs:556-564: compare byte 15 first */
static int cmp_id(const uint8_t *a, const uint8_t *b) {
  for (int i = 15; i >= 0; i--)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

/* volumes.rs:47-59 */
int po_par2_padding_width(uint32_t max_exponent) {
  if (max_exponent == 0) return 1;
  int w = 0;
  for (uint32_t n = max_exponent; n; n /= 10) w++;
  return w;
}

/* volumes.rs:61-76 */
void po_par2_volume_name(const char *stem, uint32_t start, uint32_t end, int width, char *out, size_t cap) {
  snprintf(out, cap, "%s.vol%0*u+%0*u.par2", stem, width, start, width, end - start + 1);
}

/* ---------------------------------------------------------- byte buffer -- */
typedef struct {
  uint8_t *p;
  size_t len, cap;
} Buf;

static void put(Buf *b, const void *src, size_t n) {
  if (b->len + n > b->cap) {
    size_t c = b->cap ? b->cap : 4096;
    while (c < b->len + n) c *= 2;
    b->p = realloc(b->p, c);
    b->cap = c;
  }
  memcpy(b->p + b->len, src, n);
  b->len += n;
}
static void put_u64(Buf *b, uint64_t v) {
  uint8_t t[8];
  for (int i = 0; i < 8; i++) t[i] = (uint8_t)(v >> (8 * i));
  put(b, t, 8);
}
static void put_u32(Buf *b, uint32_t v) {
  uint8_t t[4];
  for (int i = 0; i < 4; i++) t[i] = (uint8_t)(v >> (8 * i));
  put(b, t, 4);
}

/* writer.rs:83-124: magic, length, MD5(set id || type || padded body), set id,
 * type, padded body */
static void put_packet(Buf *out, const uint8_t type[16], const uint8_t set_id[16], const uint8_t *body, size_t n) {
  const size_t padded = (n + 3) & ~(size_t)3;
  Buf h = {0};
  put(&h, set_id, 16);
  put(&h, type, 16);
  put(&h, body, n);
  static const uint8_t zeros[4] = {0, 0, 0, 0};
  put(&h, zeros, padded - n);
  uint8_t md5[16];
  po_md5(h.p, h.len, md5);
  put(out, kMagic, 8);
  put_u64(out, 64 + (uint64_t)padded);
  put(out, md5, 16);
  put(out, h.p, h.len);  /* set id, type, padded body */
  free(h.p);
}

typedef struct {
  const po_file_in *in;
  uint8_t full[16], h16k[16], id[16];
  uint64_t blocks, first_block;
} FileMeta;

static int cmp_meta(const void *a, const void *b) {
  return cmp_id(((const FileMeta *)a)->id, ((const FileMeta *)b)->id);
}

int po_par2_create(const po_file_in *files, size_t nfiles, uint64_t block_size, float redundancy, int scheme,
                   const char *out_dir_stem, po_volume_out **vols_out, size_t *nvols_out) {
  *vols_out = NULL;
  *nvols_out = 0;
  if (nfiles == 0 || redundancy != redundancy || redundancy < 0.0f) return -1;
  FileMeta *meta = calloc(nfiles, sizeof *meta);
  uint64_t total = 0;
  for (size_t f = 0; f < nfiles; f++) {
    meta[f].in = &files[f];
    po_md5(files[f].data, files[f].len, meta[f].full);
    po_md5(files[f].data, files[f].len < 16384 ? files[f].len : 16384, meta[f].h16k);
    po_par2_file_id(meta[f].h16k, files[f].len, files[f].name, meta[f].id);
    total += files[f].len;
  }
  qsort(meta, nfiles, sizeof *meta, cmp_meta);  /* IDs are distinct MD5s in practice */
  if (block_size == 0) {
    block_size = po_par2_block_size(total);
  } else if (block_size < 2048 || block_size % 4) { /* with_block_size, creator.rs:185-211 */
    free(meta);
    return -1;
  }
  uint64_t n = 0;
  for (size_t f = 0; f < nfiles; f++) {
    meta[f].first_block = n;
    meta[f].blocks = (meta[f].in->len + block_size - 1) / block_size;
    n += meta[f].blocks;
  }
  const uint64_t data_bytes = n * block_size;
  const uint64_t recovery_bytes = (uint64_t)((double)data_bytes * ((double)redundancy / 100.0));
  const uint64_t r = (recovery_bytes + block_size - 1) / block_size;
  if (n > 65535 || r > 65535 || n + r > 65535 || n > 32768) {
    free(meta);
    return -2;
  }
  if (scheme == PO_SCHEME_SINGLE && r == 0) { /* creator.rs:455-457 indexes an empty list */
    free(meta);
    return -3;
  }

  /* blocks in canonical order, zero-padded; IFSC of each padded block */
  uint8_t *data = calloc(n ? n : 1, block_size);
  uint8_t *md5s = malloc((n ? n : 1) * 16);
  uint32_t *crcs = malloc((n ? n : 1) * 4);
  for (size_t f = 0; f < nfiles; f++)
    memcpy(data + meta[f].first_block * block_size, meta[f].in->data, meta[f].in->len);
  po_hash_slices(data, n, block_size, block_size, NULL, md5s, crcs);
  uint16_t *rec = malloc((r ? r : 1) * block_size);
  uint32_t *exps = malloc((r ? r : 1) * 4);
  for (uint64_t e = 0; e < r; e++) exps[e] = (uint32_t)e;
  if (r && n) po_encode((const uint16_t *)data, n, block_size / 2, exps, r, rec);
  if (r && !n) memset(rec, 0, r * block_size);

  /* main body and recovery set ID */
  Buf mainb = {0};
  put_u64(&mainb, block_size);
  put_u32(&mainb, (uint32_t)nfiles);
  for (size_t f = 0; f < nfiles; f++) put(&mainb, meta[f].id, 16);
  uint8_t set_id[16];
  po_md5(mainb.p, mainb.len, set_id);

  /* the packets every output file starts with */
  Buf head = {0};
  put_packet(&head, kTypeCreator, set_id, (const uint8_t *)po_par2_creator_text, strlen(po_par2_creator_text));
  put_packet(&head, kTypeMain, set_id, mainb.p, mainb.len);
  for (size_t f = 0; f < nfiles; f++) {
    Buf b = {0};
    put(&b, meta[f].id, 16);
    put(&b, meta[f].full, 16);
    put(&b, meta[f].h16k, 16);
    put_u64(&b, meta[f].in->len);
    put(&b, meta[f].in->name, strlen(meta[f].in->name) + 1);
    put_packet(&head, kTypeFileDesc, set_id, b.p, b.len);
    free(b.p);
  }
  for (size_t f = 0; f < nfiles; f++) {
    Buf b = {0};
    put(&b, meta[f].id, 16);
    for (uint64_t k = 0; k < meta[f].blocks; k++) {
      put(&b, md5s + 16 * (meta[f].first_block + k), 16);
      put_u32(&b, crcs[meta[f].first_block + k]);
    }
    put_packet(&head, kTypeIfsc, set_id, b.p, b.len);
    free(b.p);
  }

  /* volumes: exponent ranges and names */
  size_t nv = 0, cap = 8;
  po_volume_out *v = calloc(cap, sizeof *v);
  uint64_t *vstart = calloc(cap, sizeof *vstart), *vend = calloc(cap, sizeof *vend);
  const int width = po_par2_padding_width(r ? (uint32_t)(r - 1) : 0);
  if (scheme == PO_SCHEME_SINGLE) {
    snprintf(v[0].name, sizeof v[0].name, "%s.par2", out_dir_stem);
    vstart[0] = 0;
    vend[0] = r;
    nv = 1;
  } else {
    snprintf(v[0].name, sizeof v[0].name, "%s.par2", out_dir_stem);
    vstart[0] = vend[0] = 0;
    nv = 1;
    uint64_t e = 0, target = 1;
    while (e < r) {
      const uint64_t cnt = (r - e) < target ? (r - e) : target;
      if (nv == cap) {
        cap *= 2;
        v = realloc(v, cap * sizeof *v);
        vstart = realloc(vstart, cap * sizeof *vstart);
        vend = realloc(vend, cap * sizeof *vend);
      }
      memset(&v[nv], 0, sizeof v[nv]);
      po_par2_volume_name(out_dir_stem, (uint32_t)e, (uint32_t)(e + cnt - 1), width, v[nv].name, sizeof v[nv].name);
      vstart[nv] = e;
      vend[nv] = e + cnt;
      nv++;
      e += cnt;
      target = target * 2 < 1024 ? target * 2 : 1024;
    }
  }
  for (size_t i = 0; i < nv; i++) {
    Buf out = {0};
    put(&out, head.p, head.len);
    for (uint64_t e = vstart[i]; e < vend[i]; e++) {
      Buf b = {0};
      put_u32(&b, (uint32_t)e);
      put(&b, (const uint8_t *)rec + e * block_size, block_size);
      put_packet(&out, kTypeRecv, set_id, b.p, b.len);
      free(b.p);
    }
    v[i].bytes = out.p;
    v[i].len = out.len;
  }
  free(vstart);
  free(vend);
  free(head.p);
  free(mainb.p);
  free(rec);
  free(exps);
  free(data);
  free(md5s);
  free(crcs);
  free(meta);
  *vols_out = v;
  *nvols_out = nv;
  return 0;
}

void po_par2_free(po_volume_out *vols, size_t nvols) {
  for (size_t i = 0; i < nvols; i++) free(vols[i].bytes);
  free(vols);
}
