/*
 * par2_oracle.h -- CPU restatement of par2-rs's Reed-Solomon + slice-hash path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (libpar2hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the par2-rs 0.3.1 crate root).  Plain C99, scalar, single-threaded except
 * where noted.
 */
#ifndef PAR2_ORACLE_H
#define PAR2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- GF(2^16), poly 0x1100B, generator 2 (src/galois/core.rs:12, :49-91) ---- */
void     po_init(void);
/* Copies of the tables: log[65536] (log[0]=0), exp[131070] (doubled). */
void     po_tables(uint16_t *log_out, uint16_t *exp_out);
uint16_t po_gf_mul(uint16_t a, uint16_t b);             /* core.rs:175-189 */
int      po_gf_div(uint16_t a, uint16_t b, uint16_t *q); /* core.rs:205-230; -1 if b==0 */
uint16_t po_gf_pow(uint16_t a, uint64_t n);             /* core.rs:242-258 */

/* PAR2 input-slice constants c_k = 2^n_k (src/par2_rs.rs:64-76). */
void po_par2_constants(size_t count, uint16_t *out);

/* Coefficient matrix, stored [r][d] row-major: coeff[r*n + d] = c_d^exps[r]
 * (src/par2_rs.rs:707-719 builds c_d^r by repeated multiply; :894-900 sets
 * exponent = r).  Uses repeated multiplication when exps[r] == r for all r,
 * gf_pow otherwise (identical values, checked by tests). */
void po_coeff_matrix(const uint16_t *constants, size_t n, const uint32_t *exps,
                     size_t r, uint16_t *coeff);

/* ---- slice ops with the scalar strategy's semantics (src/galois/scalar.rs) ---- */
void po_mul_slice(uint16_t k, uint16_t *data, size_t len);                 /* :9-17  */
void po_muladd(uint16_t *dst, const uint16_t *src, size_t len, uint16_t k); /* :21-28 */
/* coeff is [ndst][coeff_stride]; only the first nsrc entries of a row are used
 * (scalar.rs:82-113, ".take(num_srcs)"). Operates on [off, off+len). */
void po_muladd_region(uint16_t *const *dsts, size_t ndst, const uint16_t *const *srcs,
                      size_t nsrc, const uint16_t *coeff, size_t coeff_stride,
                      size_t off, size_t len);

/* ---- encode: recovery[r][w] = XOR_d coeff[r][d]*data[d][w] (src/par2_rs.rs:785-902) ----
 * data: n slices of `words` u16 each, contiguous with stride `words`.
 * rec:  r slices, zero-initialised by the callee.                              */
void po_encode(const uint16_t *data, size_t n, size_t words,
               const uint32_t *exps, size_t r, uint16_t *rec);

/* ---- repair, faithful to src/par2_rs.rs:319-422 + :442-675 ----
 * missing[m], present[n-m] (global data indices), rec_exps[m] (exponents of the
 * m recovery slices used, in the caller's order).
 * Outputs of the transform: row_order[m], scale[m], elim[m*m], vand[m*(n-m)].
 * Returns 0, -1 ("Singular matrix") or -2 (m == 0). */
int po_transform(const uint16_t *constants, const size_t *missing, size_t m,
                 const size_t *present, size_t npresent, const uint32_t *rec_exps,
                 size_t *row_order, uint16_t *scale, uint16_t *elim, uint16_t *vand);

/* Chunk reconstruction with the transform above (par2_rs.rs:442-675).
 * rec_chunks[i]  : recovery slice i (caller order), `words` u16.
 * data_chunks[j] : present data slice j (same order as `present`).
 * out[i]         : rebuilt chunk for missing[i].                               */
void po_reconstruct_chunk(size_t m, size_t npresent, const size_t *row_order,
                          const uint16_t *scale, const uint16_t *elim, const uint16_t *vand,
                          const uint16_t *const *rec_chunks, const uint16_t *const *data_chunks,
                          size_t words, uint16_t *const *out);

/* Plain Gauss-Jordan inverse over GF(2^16) with row pivoting (no reference
 * counterpart as a whole; pivot rule follows par2_rs.rs:358-371).
 * a: m*m row-major input (not modified); inv: m*m output.
 * Returns 0, or -1 if singular.                                                 */
int po_gf_invert(const uint16_t *a, size_t m, uint16_t *inv);

/* ---- hashes (RFC 1321 MD5; CRC-32/IEEE reflected 0xEDB88320) ----
 * Restated from the public specifications; the reference calls md-5 0.10.6 and
 * crc32fast 1.5.0 (Cargo.lock:366-383, :128-135) at src/hash/mod.rs:22-27,
 * src/creator.rs:397-409, src/verify.rs:621,:634.                               */
void     po_md5(const uint8_t *data, size_t len, uint8_t out[16]);
uint32_t po_crc32(const uint8_t *data, size_t len);
uint32_t po_crc32_update(uint32_t crc, const uint8_t *data, size_t len); /* raw state */

/* IFSC entries for n slices: slice i is bytes [i*slice_stride, +lens[i]) of
 * `data`, zero-padded to slice_size (creator.rs:389-409, verify.rs:592-595).
 * lens == NULL means every slice is full.                                       */
void po_hash_slices(const uint8_t *data, size_t n, size_t slice_size, size_t slice_stride,
                    const size_t *lens, uint8_t *md5_out, uint32_t *crc_out);

/* ---- CPU baseline: restatement of the reference's selected x86 strategy
 * "AVX2-Shuffle" nibble-pshufb muladd (src/galois/simd/x86/avx2.rs:260-387)
 * driven with the StreamingEncoder batching (src/par2_rs.rs:88-126, :816-870),
 * `threads` worker threads (0 = all online cores).  Same result as po_encode. */
int  po_encode_avx2(const uint16_t *data, size_t n, size_t words,
                    const uint32_t *exps, size_t r, uint16_t *rec, int threads);
/* Multithreaded slice hashing (rayon par_iter over slices, verify.rs:610-648). */
void po_hash_slices_mt(const uint8_t *data, size_t n, size_t slice_size,
                       uint8_t *md5_out, uint32_t *crc_out, int threads);
int  po_have_avx2(void);

#ifdef __cplusplus
}
#endif
#endif
