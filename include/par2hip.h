/*
 * par2hip.h -- C ABI of the MI355X (gfx950) Reed-Solomon + slice-hash library.
 *
 * Drop-in boundary for par2-rs 0.3.1.  Two levels (paths relative to the
 * reference crate root):
 *   1. strategy-shaped entry points, one per method of the plugin trait
 *      `GaloisSimdStrategy` (src/galois/simd/mod.rs:24-128), so a GPU strategy
 *      can be registered in `x86::register_strategies`
 *      (src/galois/simd/x86/mod.rs:13-26);
 *   2. coarse whole-set entry points replacing the encoder
 *      (`StreamingEncoder::{process_block,flush_batch,finalize}`,
 *      src/par2_rs.rs:785-902), the repair transform and chunk loop
 *      (`compute_reconstruction_transform`, src/par2_rs.rs:319-422;
 *      `reconstruct_streaming_chunk`, src/par2_rs.rs:442-675) and the IFSC slice
 *      hashing (src/creator.rs:397-409, src/verify.rs:548-654).
 *
 * Conventions
 *   - every call returns int32_t: PAR2HIP_OK (0) or a negative PAR2HIP_E*;
 *     nothing throws or aborts across the boundary (the reference builds with
 *     panic="abort", Cargo.toml:46, so failures must be values);
 *   - one par2hip_ctx per device; a ctx owns one HIP stream and serialises
 *     calls made on it (calls from several threads on one ctx are safe);
 *   - "host" buffers are borrowed for the duration of a synchronous call;
 *     "_dev" entry points take device pointers obtained from par2hip_dev_alloc
 *     and are asynchronous on the ctx stream until par2hip_sync;
 *   - GF data are little-endian interleaved u16 words (the reference's
 *     non-shuffle2x layout, supports_shuffle2x() == false); lengths of GF
 *     operands are in words, lengths of hashed data in bytes;
 *   - coefficient matrices are row-major [dst][src].
 */
#ifndef PAR2HIP_H
#define PAR2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PAR2HIP_OK = 0,
    PAR2HIP_EINVAL = -1,    /* bad argument (sizes, null pointers, odd slice size) */
    PAR2HIP_ENODEV = -2,    /* no usable gfx950 device */
    PAR2HIP_ENOMEM = -3,    /* device allocation failed */
    PAR2HIP_ESINGULAR = -4, /* erasure matrix singular ("Singular matrix", par2_rs.rs:369) */
    PAR2HIP_EHIP = -5,      /* HIP runtime error; see par2hip_last_error */
    PAR2HIP_ERECOVERY = -6, /* fewer recovery slices than erasures (par2_rs.rs:332-338) */
    PAR2HIP_EREPAIR = -7,   /* rebuilt files failed verification; nothing was committed
                               (RepairFailed, src/repair.rs:1050-1080) */
    PAR2HIP_ECANCELLED = -8, /* the caller's cancel flag was set; no original file was
                                changed (Par2Error::Cancelled, src/error.rs:35) */
    PAR2HIP_ETIMEDOUT = -9   /* par2hip_server_wait_timeout: not done in time (no
                                counterpart in the reference) */
};

/* Strategy priority, same scale as `Priority` (src/galois/simd/mod.rs:15-21). */
enum { PAR2HIP_PRIORITY_OPTIMAL = 40 };

typedef struct par2hip_ctx par2hip_ctx;

/* ---------------------------------------------------------------- device -- */

/* GaloisSimdStrategy::is_available (simd/mod.rs:29): number of gfx950
 * devices (0 if none).  Cheap after the first call. */
int32_t par2hip_available(void);
/* GaloisSimdStrategy::name / priority (simd/mod.rs:26, :32). */
const char *par2hip_strategy_name(void);
int32_t par2hip_strategy_priority(void);

int32_t par2hip_ctx_create(int32_t device, par2hip_ctx **out);
int32_t par2hip_ctx_destroy(par2hip_ctx *ctx);
/* Text of the last error on this ctx (valid until the next call on it). */
const char *par2hip_last_error(par2hip_ctx *ctx);

/* What the ctx runs on: compute units, the CUs reserved for the MD5 side
 * stream (0 when CU masks are unavailable), and the gfx arch name. */
typedef struct {
    int32_t compute_units;
    int32_t hash_cus;
    char arch[32];
} par2hip_device_info;
int32_t par2hip_get_device_info(par2hip_ctx *ctx, par2hip_device_info *out);

/* Device memory owned by the ctx's device (serving keeps sets resident). */
int32_t par2hip_dev_alloc(par2hip_ctx *ctx, size_t bytes, void **dptr);
int32_t par2hip_dev_free(par2hip_ctx *ctx, void *dptr);
int32_t par2hip_h2d(par2hip_ctx *ctx, void *dst, const void *src, size_t bytes);
int32_t par2hip_d2h(par2hip_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy, asynchronous on the ctx stream. */
int32_t par2hip_d2d(par2hip_ctx *ctx, void *dst, const void *src, size_t bytes);
int32_t par2hip_sync(par2hip_ctx *ctx);

/* --------------------------------------------- strategy-shaped (host) -- */

/* GaloisSimdStrategy::mul_slice (simd/mod.rs:38): data[i] = data[i]*scalar;
 * scalar 0 zeroes, 1 is a no-op (src/galois/scalar.rs:9-17). */
int32_t par2hip_gf_mul_slice(par2hip_ctx *ctx, uint16_t scalar, uint16_t *data, size_t len);

/* GaloisSimdStrategy::muladd (simd/mod.rs:44): dst[i] ^= src[i]*scalar for
 * i < len; scalar 0 is a no-op (src/galois/scalar.rs:21-28). */
int32_t par2hip_gf_muladd(par2hip_ctx *ctx, uint16_t *dst, const uint16_t *src, size_t len,
                          uint16_t scalar);

/* GaloisSimdStrategy::muladd_region (simd/mod.rs:50-57): for every dst i and
 * src j < min(coeff_lens[i], nsrc):
 *   dst_i[off..off+len) ^= coeff_rows[i][j] * src_j[off..off+len)
 * (src/galois/scalar.rs:82-113; a coefficient row shorter than the source
 * list applies only its own entries, as `.enumerate().take(num_srcs)` does at
 * scalar.rs:100).  coeff_rows[i] points at coeff_lens[i] values; coeff_lens
 * NULL means every row holds at least nsrc. */
int32_t par2hip_gf_muladd_region(par2hip_ctx *ctx, uint16_t *const *dsts, size_t ndst,
                                 const uint16_t *const *srcs, size_t nsrc,
                                 const uint16_t *const *coeff_rows, const size_t *coeff_lens,
                                 size_t region_offset, size_t region_size);

/* ------------------------------------------------------------ encode -- */

/* Replaces StreamingEncoder (src/par2_rs.rs:687-902) for a whole set.
 * slices:   n input slices of slice_size bytes each, contiguous (slice d at
 *           slices + d*slice_size), already zero-padded (creator.rs:393-395).
 * exps:     r recovery exponents (the reference uses 0..r-1, par2_rs.rs:894-900).
 * recovery: r*slice_size bytes out; recovery e = XOR_d c_d^exps[e] * slice_d,
 *           c_d the PAR2 constants (par2_rs.rs:64-76).
 * md5_out / crc_out (nullable): IFSC MD5 (16 B) and CRC32 of every input
 *           slice (creator.rs:397-409).
 * slice_size must be even and non-zero; n <= 32768 (distinct constants);
 * with MD5 requested, slice_size < 2^37 (128 GiB: the MD5 kernels count
 * 64-byte blocks in 32 bits), else PAR2HIP_EINVAL before any launch. */
int32_t par2hip_encode(par2hip_ctx *ctx, const uint8_t *slices, size_t n, size_t slice_size,
                       const uint32_t *exps, size_t r, uint8_t *recovery, uint8_t *md5_out,
                       uint32_t *crc_out);

/* Same, device-resident and asynchronous on the ctx stream. d_md5 / d_crc
 * may be NULL. */
int32_t par2hip_encode_dev(par2hip_ctx *ctx, const void *d_slices, size_t n, size_t slice_size,
                           const uint32_t *exps, size_t r, void *d_recovery, void *d_md5,
                           void *d_crc);

/* par2hip_encode_dev without any allocation during the call.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (work captured into a HIP graph, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and asynchrony as par2hip_encode_dev
 * (recovery slices bit-identical; the contraction is always the LDS-table
 * form).  Its working memory is one 64 MiB device array per device, held
 * until the process ends: an encode needs 256-byte-rounded n*4 + r*4 + n*8 +
 * r*8 + r*n*2 + ceil(r/8)*n*1024 bytes of it (the 1 GiB / 1024-slice set with
 * r = 102 needs ~14 MB), and PAR2HIP_EINVAL is returned, before anything is
 * queued, when that exceeds 64 MiB.  Every call on a device shares that
 * array: calls that overlap on one device (through different contexts or
 * threads) race; calls queued on one context run in order. */
int32_t par2hip_encode_dev_noalloc(par2hip_ctx *ctx, const void *d_slices, size_t n, size_t slice_size,
                                   const uint32_t *exps, size_t r, void *d_recovery, void *d_md5,
                                   void *d_crc);

/* ------------------------------------------------------------ repair -- */

/* Gauss-Jordan inverse of an m x m GF(2^16) matrix on the device
 * (replaces the O(m^3) elimination of par2_rs.rs:358-412).  Host buffers.
 * Returns PAR2HIP_ESINGULAR if no pivot exists in some column. */
int32_t par2hip_gf_invert(par2hip_ctx *ctx, const uint16_t *a, size_t m, uint16_t *inv);

/* Rebuild the m missing input slices of an n-slice set (replaces
 * compute_reconstruction_transform + reconstruct_streaming_chunk over all
 * chunks, repair.rs:848-970).
 * slices:     n*slice_size bytes; the missing slices' contents are ignored.
 * missing:    m distinct global slice indices (any order).
 * recovery:   m recovery slices (m*slice_size bytes) with exponents rec_exps.
 * out:        m*slice_size bytes, row i = rebuilt slice missing[i].
 * Errors: PAR2HIP_ESINGULAR, PAR2HIP_EINVAL.                                  */
int32_t par2hip_reconstruct(par2hip_ctx *ctx, const uint8_t *slices, size_t n, size_t slice_size,
                            const uint32_t *missing, size_t m, const uint8_t *recovery,
                            const uint32_t *rec_exps, uint8_t *out);

/* Device-resident form.  The call waits for the (short) inverse so it can
 * report PAR2HIP_ESINGULAR, then queues the rebuild on the ctx stream
 * (complete at par2hip_sync).  With out_in_place non-zero the rebuilt slices
 * are written into their rows of d_slices and d_out is ignored. */
int32_t par2hip_reconstruct_dev(par2hip_ctx *ctx, void *d_slices, size_t n, size_t slice_size,
                                const uint32_t *missing, size_t m, const void *d_recovery,
                                const uint32_t *rec_exps, void *d_out, int32_t out_in_place);

/* Verify a device-resident set against its IFSC checksums and rebuild the
 * damaged slices (replaces verify_file_blocks, src/verify.rs:548-654, the
 * damaged-slice collection and recovery choice of src/repair.rs:536-604, and
 * the transform + chunk loop, repair.rs:848-970).
 * d_slices:      n zero-padded slices of slice_size bytes.
 * expected_md5 / expected_crc: host arrays, the IFSC entries (n*16 B, n u32).
 * d_recovery:    nrec recovery slices (contiguous), exponents rec_exps[nrec].
 * d_out:         NULL = rebuild in place in d_slices; otherwise room for
 *                min(n, nrec) slices, row k = rebuilt damaged_out[k].
 * damaged_out:   host, capacity n: the damaged indices, ascending
 *                (damaged = CRC32 or MD5 mismatch, verify.rs:621-645);
 *                *ndamaged receives their count.
 * The ndamaged lowest recovery exponents are used (repair.rs:584-604).
 * CRC32 picks the erasures first and the rebuild starts while the MD5
 * confirmation still runs; a slice whose CRC matches but MD5 does not is
 * added and the rebuild redone.  Errors: PAR2HIP_ERECOVERY (more damaged
 * slices than recovery slices; damaged_out is still filled),
 * PAR2HIP_ESINGULAR. */
int32_t par2hip_repair_set_dev(par2hip_ctx *ctx, void *d_slices, size_t n, size_t slice_size,
                               const uint8_t *expected_md5, const uint32_t *expected_crc,
                               const void *d_recovery, const uint32_t *rec_exps, size_t nrec,
                               void *d_out, uint32_t *damaged_out, size_t *ndamaged);

/* ------------------------------------------------------------- batched -- */

/* One job of a batch: a whole PAR2 set to encode, or one to repair. */
enum { PAR2HIP_JOB_ENCODE = 0, PAR2HIP_JOB_REPAIR = 1 };
typedef struct {
    int32_t kind;            /* PAR2HIP_JOB_ENCODE or PAR2HIP_JOB_REPAIR */
    int32_t reserved;
    void *d_slices;          /* n zero-padded slices of slice_size bytes (device) */
    uint64_t n;
    uint64_t slice_size;     /* even; any multiple of 2 */
    const uint32_t *exps;    /* host. encode: the count recovery exponents;
                                repair: exponents of the count recovery rows */
    uint64_t count;          /* encode: recovery slices out; repair: erasures */
    void *d_recovery;        /* encode: out, count rows; repair: in, count rows */
    const uint32_t *missing; /* repair: count erased slice indices (host) */
    void *d_out;             /* repair: count rebuilt rows out, row k = missing[k] */
    void *d_md5;             /* nullable: IFSC MD5 of every input slice (n*16 B) */
    void *d_crc;             /* nullable: CRC32 of every input slice (n u32) */
} par2hip_job;

/* Run many jobs as one batch (no counterpart in the reference, which handles
 * one set per Par2Creator / Par2Repairer): coefficients, inverses (closed form
 * for consecutive exponents), tables and hashes of all jobs in one launch each
 * and every contraction in one persistent launch over a cost-sorted work list.
 * Asynchronous on the ctx stream; complete at par2hip_sync.  Repair jobs with
 * non-consecutive exponents take the Gauss-Jordan path after the batch. */
int32_t par2hip_batch_run_dev(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs);

/* A planned batch (no counterpart in the reference): the host planning,
 * parameter upload and working memory of par2hip_batch_run_dev done once, so
 * the same jobs (same device buffers; their contents may change) can be run
 * many times.  With use_graph the launches are captured into one HIP graph
 * and each launch replays it.  Every repair job needs consecutive recovery
 * exponents (PAR2HIP_EINVAL otherwise).  Launches are asynchronous on the ctx
 * stream; destroy waits for them. */
typedef struct par2hip_batch_plan par2hip_batch_plan;
int32_t par2hip_batch_plan_create(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs,
                                  int32_t use_graph, par2hip_batch_plan **out);
/* With tile-shape autotuning on (par2hip_set_autotune(ctx, 1)) and no forced
 * shape, par2hip_batch_plan_create plans the batch under each tile-shape
 * policy -- 0: 8 words per thread and one 8-row group per item for at most 8
 * outputs, two above; 1: the tuned table's shape for each task's class (the
 * single-set tuner's); 2: one group everywhere; 3: 4 words per thread for
 * slices of at most 64 Ki words -- times two launches of each on the batch
 * itself and keeps the fastest (policy 0 unless another is at least 2%
 * faster).  No counterpart in the reference.  The timings are readable: */
typedef struct {
    int32_t policy;       /* the policy the plan uses */
    int32_t policies;     /* policies timed (0 when the plan was not tuned) */
    double ms[4];         /* mean launch time of each policy */
} par2hip_batch_plan_tuning;
int32_t par2hip_batch_plan_tuning_info(par2hip_batch_plan *plan, par2hip_batch_plan_tuning *out);
int32_t par2hip_batch_plan_launch(par2hip_batch_plan *plan);
int32_t par2hip_batch_plan_destroy(par2hip_batch_plan *plan);

/* ------------------------------------------------------------- serving -- */

/* A request loop with continuous batching (no counterpart in the reference,
 * which runs one Par2Creator::create / Par2Repairer::repair per call:
 * src/creator.rs:261-531, src/repair.rs:736-1125).  Requests are admitted in
 * arrival order while their HBM footprint (input slices + outputs + working
 * memory) fits hbm_budget; `lanes` worker threads, each with its own context,
 * each take everything admitted so far (at most max_batch requests) as one
 * batch, so requests that arrive while batches run join the next one and the
 * batches of different lanes overlap on the device.  Input slices are
 * device-resident and caller-owned (they must stay valid until the request is
 * done); the server allocates the outputs from a pool of hbm_budget bytes and
 * keeps them until par2hip_server_release. */
typedef struct par2hip_server par2hip_server;
typedef struct {
    par2hip_job job;       /* inputs; d_recovery (encode), d_out, d_md5 and d_crc
                              are ignored and allocated by the server */
    int32_t hashes;        /* 1: also MD5/CRC32 of every input slice */
    int32_t reserved;
    double arrival_ms;     /* server clock (par2hip_server_now_ms); 0 = submit time */
} par2hip_request;
typedef struct {
    int32_t status;        /* PAR2HIP_OK or the error of this request alone */
    int32_t batch_size;    /* requests in the batch that ran this one */
    void *d_out;           /* encode: count recovery slices; repair: count rebuilt slices */
    void *d_md5;           /* n * 16 bytes, or NULL */
    void *d_crc;           /* n u32, or NULL */
    double t_arrival_ms, t_submit_ms, t_admit_ms, t_start_ms, t_done_ms;
} par2hip_result;
typedef struct {
    uint64_t submitted, completed, failed, batches, max_batch_seen;
    uint64_t charged, peak_charged;  /* bytes of the budget held now / at most */
    uint64_t pending, admitted;      /* requests waiting for memory / for a lane */
    uint64_t arena_fallbacks;        /* batches whose working memory did not come from the pool */
} par2hip_server_stats;
int32_t par2hip_server_create(int32_t device, uint64_t hbm_budget, int32_t lanes, int32_t max_batch,
                              par2hip_server **out);
/* Finishes the admitted requests, drops the never-admitted ones, frees all. */
int32_t par2hip_server_destroy(par2hip_server *srv);
double par2hip_server_now_ms(par2hip_server *srv);
/* Non-blocking.  PAR2HIP_EINVAL for a malformed job, PAR2HIP_ENOMEM for one
 * whose footprint exceeds the whole budget. */
int32_t par2hip_server_submit(par2hip_server *srv, const par2hip_request *req, uint64_t *ticket);
/* Blocks until the request is done; its outputs stay valid until release.
 * Admission is strict FIFO: a request whose footprint does not fit the free
 * budget holds back every request submitted after it until earlier requests
 * are released.  A client that waits on such a request while it still holds
 * unreleased results it needs to free the budget therefore waits forever;
 * par2hip_server_wait_timeout bounds that wait. */
int32_t par2hip_server_wait(par2hip_server *srv, uint64_t ticket, par2hip_result *res);
/* par2hip_server_wait with a limit: PAR2HIP_ETIMEDOUT (res untouched, the
 * request keeps its place) when the request is not done within timeout_ms
 * milliseconds; 0 polls; an infinite value waits as par2hip_server_wait. */
int32_t par2hip_server_wait_timeout(par2hip_server *srv, uint64_t ticket, double timeout_ms,
                                    par2hip_result *res);
int32_t par2hip_server_release(par2hip_server *srv, uint64_t ticket);
int32_t par2hip_server_get_stats(par2hip_server *srv, par2hip_server_stats *st);

/* ---------------------------------------------------------- .par2 files -- */

/* Par2Creator::create (src/creator.rs:261-531) with the slice hashes,
 * recovery slices and recovery packet digests on the device: writes
 * <out_stem>.par2 and, for the exponential scheme, the recovery volumes
 * <out_stem>.volS+C.par2 (src/volumes.rs:90-208), byte for byte as the
 * reference lays them out (packets: src/writer.rs:83-264; creator text
 * src/creator.rs:496).  files: the input files' names (as stored in the
 * FileDesc packets) and contents; block_size 0 = the reference's automatic
 * size (creator.rs:29-67), else >= 2048 and a multiple of 4;
 * redundancy_percent as the reference's f32.  outs receives the written
 * files (main .par2 first); *nouts the number needed (PAR2HIP_EINVAL when
 * max_outs is smaller, before anything is written). */
typedef struct {
    const char *name;
    const void *data;
    uint64_t len;
} par2hip_input_file;
typedef struct {
    char path[1024];
    uint64_t bytes;
} par2hip_output_file;
enum { PAR2HIP_VOLUMES_SINGLE = 0, PAR2HIP_VOLUMES_EXPONENTIAL = 1 };
int32_t par2hip_create_par2(par2hip_ctx *ctx, const par2hip_input_file *files, size_t nfiles,
                            uint64_t block_size, float redundancy_percent, int32_t scheme,
                            const char *out_stem, par2hip_output_file *outs, size_t max_outs,
                            size_t *nouts);

/* What a .par2 index and the volumes beside it hold (Par2Info::load,
 * src/info.rs; parser src/parser.rs:181-547).  Needs no device. */
typedef struct {
    uint64_t block_size;
    uint64_t blocks;            /* input blocks of the described files */
    uint32_t files;             /* described files */
    uint32_t recovery_slices;   /* distinct recovery exponents found */
    uint32_t min_exponent, max_exponent;
    uint8_t set_id[16];
} par2hip_set_info;
int32_t par2hip_par2_info(const char *par2_path, par2hip_set_info *info);

/* Par2Repairer: load par2_path and every .par2 beside it with the same
 * recovery set ID (src/parser.rs:181-547), verify the protected files in
 * base_dir (NULL: the index's directory) with IFSC block checks on the device
 * (src/verify.rs:87-330; a missing file or a size mismatch makes all of its
 * blocks damaged), rebuild the damaged blocks on the device from the lowest
 * exponents, write each repaired file to <file>.par2tmp seeded with the
 * original, verify every temp on the device and rename them over the
 * originals only when all verify (src/repair.rs:445-1125).  verify_only: stop
 * after verification.  Errors: PAR2HIP_ERECOVERY (nothing written),
 * PAR2HIP_ESINGULAR, PAR2HIP_EREPAIR (temps removed, originals untouched).
 * The context keeps a pinned host staging buffer of the set's block bytes
 * (and device buffers of the same size) for later calls until it is
 * destroyed. */
typedef struct {
    uint64_t block_size, blocks, blocks_damaged;
    uint32_t files, files_damaged, files_missing;
    uint32_t recovery_slices;
    int32_t repaired;           /* 1 when damaged files were rewritten */
    int32_t reserved;
} par2hip_repair_report;
int32_t par2hip_repair_files(par2hip_ctx *ctx, const char *par2_path, const char *base_dir,
                             int32_t verify_only, par2hip_repair_report *report);
/* Par2Repairer::repair_with_callbacks_cancellable (src/lib.rs:196-345):
 * par2hip_repair_files with a cancel flag, read with relaxed atomic loads at
 * the reference's safe points: before loading, before and after
 * verification, after the rebuild, after the temps are written and before
 * the first rename (src/repair.rs:808-813, :972-977, :1043-1067).  cancel
 * points at one byte, 0 or 1 (the layout of a Rust AtomicBool); NULL never
 * cancels.  A cancelled call returns PAR2HIP_ECANCELLED, removes every
 * .par2tmp it wrote and leaves every original byte-identical; once the
 * renames have started the repair is committed and reported as done. */
int32_t par2hip_repair_files_cancellable(par2hip_ctx *ctx, const char *par2_path, const char *base_dir,
                                         int32_t verify_only, const uint8_t *cancel,
                                         par2hip_repair_report *report);

/* -------------------------------------------------------------- hashes -- */

/* IFSC MD5 + CRC32 of n slices, each zero-padded to slice_size
 * (creator.rs:389-409; verify.rs:592-595).  Slice i is read from
 * slices + i*slice_stride for lens[i] bytes (lens NULL: slice_size each).
 * md5_out: n*16 bytes (nullable); crc_out: n u32 (nullable).  With md5_out,
 * slice_size < 2^37 (128 GiB), else PAR2HIP_EINVAL before any launch; the
 * same limit holds for par2hip_repair_set_dev and for hashed batch jobs. */
int32_t par2hip_hash_slices(par2hip_ctx *ctx, const uint8_t *slices, size_t n, size_t slice_size,
                            size_t slice_stride, const uint64_t *lens, uint8_t *md5_out,
                            uint32_t *crc_out);
int32_t par2hip_hash_slices_dev(par2hip_ctx *ctx, const void *d_slices, size_t n,
                                size_t slice_size, size_t slice_stride, const uint64_t *lens,
                                void *d_md5, void *d_crc);

/* ------------------------------------------------------------- tuning -- */

/* GF contraction kernel selection: 0 = automatic (the LDS tables, except an
 * encode whose MD5 chains run on the reserved CUs: there the FP4 form on as
 * few CUs as finish under the chains, which keeps the chip's clock up for
 * them), 1 = LDS nibble tables, 2 = bit-sliced MFMA with FP4 operands, 3 =
 * the bit-sliced MFMA with int8 operands.  No counterpart in the reference.
 * Returns the previous value. */
int32_t par2hip_set_gf_kernel(par2hip_ctx *ctx, int32_t which);

/* Tile-shape autotuning of the LDS contraction (no counterpart in the
 * reference).  mode 0: a fixed heuristic picks the tile shape (words per
 * thread, output rows per workgroup, sources per LDS chunk); mode 1: the first
 * overwriting contraction of each shape class (rows, sources, words,
 * alignment) on this ctx times every candidate shape on that problem and the
 * fastest is kept for the process.  Returns the previous mode. */
int32_t par2hip_set_autotune(par2hip_ctx *ctx, int32_t mode);
/* Use tile shape (v words per thread, g groups of 8 rows per workgroup, d
 * sources per LDS chunk) for every LDS contraction on this ctx whose
 * alignment allows v (for tests and manual tuning); (0, 0, 0) clears it.
 * PAR2HIP_EINVAL for a shape that is not among the candidates. */
int32_t par2hip_force_tile_shape(par2hip_ctx *ctx, int32_t v, int32_t g, int32_t d);
/* The process-wide table of tuned shapes, as text (one "key v g d" line
 * each): saved to a file, or loaded from one (entries add to and replace
 * those in the table; *loaded receives the number read; a malformed line or
 * an unknown shape gives PAR2HIP_EINVAL, keeping the lines before it). */
int32_t par2hip_autotune_save(const char *path);
int32_t par2hip_autotune_load(const char *path, size_t *loaded);

/* Per-kernel device timings, recorded with HIP events on the stream each
 * kernel runs on, while enabled. */
typedef struct {
    char name[48];
    uint64_t launches;
    double total_ms;
} par2hip_kernel_stat;
int32_t par2hip_profile_enable(par2hip_ctx *ctx, int32_t on);
/* Copies up to max entries; *count receives the number available.  Reading
 * synchronises the ctx stream and resets the counters. */
int32_t par2hip_profile_read(par2hip_ctx *ctx, par2hip_kernel_stat *out, size_t max,
                             size_t *count);

#ifdef __cplusplus
}
#endif
#endif /* PAR2HIP_H */
