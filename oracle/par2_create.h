/*
 * par2_create.h -- CPU oracle (test infrastructure only): the reference's PAR2
 * creation restated in memory (src/creator.rs:261-531, src/writer.rs,
 * src/volumes.rs).  Only the tests use it.
 */
#ifndef PAR2_CREATE_H
#define PAR2_CREATE_H
#include <stddef.h>
#include <stdint.h>

typedef struct {
  const char *name;     /* file name as stored in the FileDesc packet */
  const uint8_t *data;  /* file contents */
  uint64_t len;
} po_file_in;

typedef struct {
  char name[1024];      /* output path: <stem>.par2 or <stem>.volS+C.par2 */
  uint8_t *bytes;
  uint64_t len;
} po_volume_out;

enum { PO_SCHEME_SINGLE = 0, PO_SCHEME_EXPONENTIAL = 1 };

extern const char *const po_par2_creator_text;

/* calculate_block_size (creator.rs:29-67) */
uint64_t po_par2_block_size(uint64_t total_size);
/* File ID = MD5(hash_16k || length LE u64 || name) (writer.rs:56-63) */
void po_par2_file_id(const uint8_t hash16k[16], uint64_t length, const char *name, uint8_t out[16]);
/* volumes.rs:47-76 */
int po_par2_padding_width(uint32_t max_exponent);
void po_par2_volume_name(const char *stem, uint32_t start, uint32_t end, int width, char *out, size_t cap);

/* Par2Creator::create (creator.rs:261-531) on in-memory files.  block_size 0
 * = automatic, else at least 2048 and a multiple of 4 (creator.rs:185-211); redundancy is the f32 percentage; out_stem is the output path
 * without ".par2".  Returns 0 and the output files (first the main .par2), or
 * -1 bad argument, -2 too many blocks (data + recovery > 65535, or more than
 * 32768 data blocks), -3 single scheme without recovery blocks. */
int po_par2_create(const po_file_in *files, size_t nfiles, uint64_t block_size, float redundancy, int scheme,
                   const char *out_stem, po_volume_out **vols, size_t *nvols);
void po_par2_free(po_volume_out *vols, size_t nvols);

#endif
