// common.hpp -- shared internals of libpar2hip: context, errors, GF(2^16)
// host tables and the launch helpers every kernel file uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/par2hip.h"

namespace par2hip {

// GF(2^16) with the PAR2 polynomial x^16+x^12+x^3+x+1 and generator 2
// (src/galois/core.rs:12, :49-91).
constexpr uint32_t kPoly = 0x1100Bu;
constexpr uint32_t kOrder = 65535u;

struct HostField {
  uint16_t log[65536];
  uint16_t exp[2 * kOrder];  // doubled: exp[a+b] needs no reduction
  HostField();
  uint16_t mul(uint16_t a, uint16_t b) const {
    return (a && b) ? exp[uint32_t(log[a]) + log[b]] : 0;
  }
};
const HostField &field();

// Exponent n_d of the d-th PAR2 input constant c_d = 2^n_d: the d-th n >= 1
// with n mod 3, 5, 17, 257 all non-zero (src/par2_rs.rs:64-76).  Because the
// generator is 2, log(c_d) = n_d, so coefficients are exp[(n_d*e) mod 65535].
std::vector<uint32_t> constant_logs(size_t count);

struct KernelStat {
  std::string name;
  uint64_t launches = 0;
  double total_ms = 0;
};

struct PendingTiming {
  int stat;
  hipEvent_t start, stop;
};

// A growable pair of (pinned host, device) buffers used to ship small
// per-call parameter blocks (pointer tables, exponent lists).
struct ParamArena {
  uint8_t *host = nullptr;
  uint8_t *dev = nullptr;
  size_t cap = 0;
  size_t used = 0;
  hipEvent_t copied = nullptr;  // the last H2D out of `host` has finished
  uint8_t *target = nullptr;    // device block of the current call (dev unless a plan owns one)
};

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace par2hip

struct par2hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;       // everything is ordered on this stream
  hipStream_t side = nullptr;         // MD5 chains overlap here (a few reserved CUs), joined per call
  hipStream_t rest = nullptr;         // the other CUs, for work that overlaps the side stream
  hipStream_t cur = nullptr;          // stream the current call's main work goes to
  bool masked = false;                // side/rest carry complementary CU masks
  int hash_cus = 0;                   // CUs of the side stream's mask
  hipStream_t wide = nullptr;         // unmasked helper stream (hash loads too big for the side CUs)
  // narrower complements of the side stream (narrow_cus[i] CUs each), for a
  // contraction that only has to finish under the MD5 chains and should draw
  // as little power as that allows; rest_cur is the one in use until join
  hipStream_t narrow[3] = {nullptr, nullptr, nullptr};
  int narrow_cus[3] = {0, 0, 0};
  hipStream_t rest_cur = nullptr;
  char arch[32] = {0};
  hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;
  std::recursive_mutex mu;
  std::string err;
  uint16_t *d_exp = nullptr;          // device copy of the doubled exp table
  uint16_t *d_log = nullptr;          // device copy of the log table
  uint32_t *d_crc_tab = nullptr;      // CRC-32 slice-by-4 tables T0..T3, then the row-skip tables Z0..Z3 (8 x 256)
  int *d_status = nullptr;            // device status word (singular flag)
  par2hip::ParamArena arena;
  par2hip::DevBuf scratch[12];        // grow-only device scratch, by role
  uint8_t *host_stage = nullptr;      // grow-only pinned host staging for file I/O
  size_t host_stage_bytes = 0;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};  // the staging halves' last copies (h2d_staged / d2h_staged)
  int gf_kernel = 0;                  // 0 auto, 1 LDS tables, 2 MFMA FP4, 3 MFMA int8
  int num_cus = 256;                  // compute units of the device
  bool profile = false;
  bool autotune = false;  // time candidate tile shapes on first use of a shape class
  int forced_shape = -1;  // tile shape used by every LDS contraction (-1: none)
  void *noalloc_ws = nullptr;  // this device's par2hip_encode_dev_noalloc workspace
  std::vector<par2hip::KernelStat> stats;
  std::vector<par2hip::PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
};

namespace par2hip {

enum ScratchRole {
  kScratchCoeff = 0,   // coefficient matrix
  kScratchTables = 1,  // per-coefficient lookup tables / MFMA operands
  kScratchGj = 2,      // Gauss-Jordan workspace
  kScratchIn = 3,      // host-call staging: inputs
  kScratchOut = 4,     // host-call staging: outputs
  kScratchHash = 5,    // host-call staging: digests
  kScratchAux = 6,
  kScratchAux2 = 7,
  kScratchVand = 8,    // repair: V rows c_present^e (fold sources)
  kScratchHashD = 9,   // repair: digests of the verify pass
  kScratchBatch = 10,  // batched runs: all per-set workspaces in one arena
  kScratchGjWork = 11, // blocked Gauss-Jordan: panel, pivots, pointer tables
};

int32_t fail(par2hip_ctx *ctx, int32_t code, const std::string &msg);
int32_t hip_fail(par2hip_ctx *ctx, hipError_t e, const char *what);
#define P2H_CHECK(ctx, expr)                                     \
  do {                                                           \
    hipError_t e_ = (expr);                                      \
    if (e_ != hipSuccess) return ::par2hip::hip_fail(ctx, e_, #expr); \
  } while (0)

// Device scratch of at least `bytes` for a role (contents not preserved on growth).
int32_t scratch(par2hip_ctx *ctx, int role, size_t bytes, void **out);
// Pinned host buffer of at least `bytes` (grow-only, kept by the context):
// files are read into it and copied to the device at pinned-memory rates.
int32_t host_stage(par2hip_ctx *ctx, size_t bytes, uint8_t **out);
// par2hip_ctx_create, optionally without the narrow CU-masked streams
int32_t ctx_create(int32_t device, bool narrow_streams, par2hip_ctx **out);
// memcpy of many ranges in 8 MiB pieces over up to max_threads host threads
struct CopyJob {
  uint8_t *dst;
  const uint8_t *src;
  uint64_t n;
};
void copy_parallel(const std::vector<CopyJob> &jobs, size_t max_threads);
// Copies between caller (pageable) memory and the device through the pinned
// staging buffer, in chunks: host threads fill or drain one half while the
// other half's DMA runs (on ctx->stream).  h2d_staged returns once the
// caller's bytes have been read; d2h_staged once they have been written.
int32_t h2d_staged(par2hip_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int32_t d2h_staged(par2hip_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);

// Parameter blocks: begin() waits for the previous upload, alloc() carves
// 256-byte aligned pieces (returns host and device addresses), commit()
// uploads everything on the ctx stream.
// One parameter block per API call: host staging in `arena.host`, device
// copy in `target` (a plan-owned block of at least `bytes`) or the ctx's own.
int32_t params_begin(par2hip_ctx *ctx, size_t bytes, uint8_t *target = nullptr);
void *params_alloc(par2hip_ctx *ctx, size_t bytes, void **dev);
int32_t params_commit(par2hip_ctx *ctx);

// Kernel timing (only when profiling is enabled).
int timing_start(par2hip_ctx *ctx, const char *name, hipStream_t s, hipEvent_t *stop);
void timing_stop(par2hip_ctx *ctx, int slot, hipStream_t s, hipEvent_t stop);

struct ScopedTiming {
  par2hip_ctx *ctx;
  hipStream_t s;
  int slot;
  hipEvent_t stop = nullptr;
  ScopedTiming(par2hip_ctx *c, const char *name, hipStream_t st) : ctx(c), s(st) {
    slot = timing_start(c, name, st, &stop);
  }
  ~ScopedTiming() { timing_stop(ctx, slot, s, stop); }
};

// ---- kernels (device code in gf_kernels.hip, hash_kernels.hip, gj_kernels.hip)

// coeff[r][d] = exp[(src_log[d] * exps[r]) mod 65535]; ld = nsrc.
int32_t launch_coeff_from_logs(par2hip_ctx *ctx, hipStream_t s, uint16_t *coeff,
                               const uint32_t *d_src_log, const uint32_t *d_exps, int nrow,
                               int nsrc);

// GF contraction: dst_i (=|^=) XOR_j coeff[i][j] * src_j over `words` words.
struct GfProblem {
  const uint16_t *const *srcs;  // device array of nsrc device pointers
  uint16_t *const *dsts;        // device array of ndst device pointers
  const uint16_t *coeff;        // device [ndst][ldc]
  int nsrc, ndst;
  int ldc = 0;                  // row stride of coeff in entries (0: nsrc)
  size_t words;
  bool accumulate;
  int align_words;              // every pointer and `words` stride aligned to this many words
  const char *tag = nullptr;    // kernel name for timing (default: gf_muladd_<variant>)
  int kernel = 0;               // 1 LDS tables, 2 / 3 MFMA FP4 / int8; 0: the context's choice
  bool beside_chains = false;   // runs beside MD5 chains on the other CUs (FP4 tile shape choice)
};
// Rate of the bit-sliced FP4 contraction per CU (binary multiply-adds per
// second) measured on the 1 GiB / R=102 encode: 1.40e13 over 5.02 ms on 256
// CUs alone and over 8.04 ms on 160 CUs beside the MD5 chains.
constexpr double kFp4MacPerCuSec = 1.09e13;
// Estimated time of n MD5 chains of `bytes` each (zero padding and length
// block included) on `simds` SIMDs, one wave of 64 chains per SIMD at a time,
// from the issue slots per 64-byte block of the form that runs (helper: the
// helper-wave form), at the 2.4 GHz peak shader clock.
double md5_chains_ms(size_t n, uint64_t bytes, int simds, bool helper);
// Whether launch_md5 runs the helper-wave form for these arguments.
bool md5_helper_form(const par2hip_ctx *ctx, hipStream_t s, const void *d_data, size_t n, size_t slice_size,
                     size_t stride, bool lens);
int32_t run_gf(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p);
// Largest per-coefficient operand block (tables) one contraction launch gets;
// larger problems are split by output rows.
constexpr long kGfChunkBytes = 1L << 30;
// The process-wide table of tuned tile shapes as text (one "key v g d" line each).
int32_t tune_table_save(const char *path);
// Index of tile shape (v, g, d) among the candidates, or -1.
int tile_shape_index(int v, int g, int d);
int32_t tune_table_load(const char *path, size_t *loaded);

// ---- variant T nibble tables.  One 16-byte entry per (group of 8 output
// rows g, source d, nibble position i, nibble value v), entry index
// (g * nsrc + d) * 64 + 16 * i + v; it packs the 8 rows' products
// c_{8g+j,d} * (v << 4i) as u16 j (little-endian, so u32 j/2, half j%2).
// A workgroup covers G groups = 8G output rows (G = 1, 2 or 4).
constexpr int kTabRows = 8;
__host__ __device__ inline int gf_tab_groups(int ndst) { return (ndst + kTabRows - 1) / kTabRows; }
__host__ __device__ inline long gf_tab_entries(int ndst, int nsrc) { return (long)gf_tab_groups(ndst) * nsrc * 64; }
__host__ __device__ inline int gf_rowblks(int ndst, int groups) { return (gf_tab_groups(ndst) + groups - 1) / groups; }
// Groups per workgroup of a batched task: one for at most 8 rows (no padded
// group), two otherwise.
__host__ __device__ inline int gf_batch_groups(int ndst) { return ndst <= kTabRows ? 1 : 2; }

// ---- batched contraction (many sets in one persistent launch)
struct GfTaskDev {
  const uint16_t *const *srcs;
  uint16_t *const *dsts;
  const uint16_t *coeff;  // [ndst][ldc]
  uint64_t *tabs;         // nibble tables of this task, gf_tab_entries(ndst, nsrc) 16-byte entries
  long words;
  int nsrc, ndst, ldc, accumulate;
  int align;   // byte alignment (2..16) shared by every source and destination row
  int groups;  // 8-row groups per workgroup (1 or 2)
  int v;       // words per thread (8, 4, 2 or 1): a work item covers kTpb * v words
  long pair0;  // first (8-row group, source) pair of this task in the batch's table numbering
};
struct GfItem {
  int task, tile, rowblk, pad_;
};
// Tile shape (words per thread v, 8-row groups g; 16 sources per LDS chunk)
// of one batched task under a policy: a forced shape always wins.  The persistent batch kernel has the shapes with
// v in {8, 4} and g <= 2 (its LDS holds 16 sources x 2 groups; more variants
// inlined into the one persistent kernel spill registers): a tuned g = 4 or
// 32-source shape maps to the same v with g = 2, a narrower v to the default.
// Shape policies a batch plan is tuned over (par2hip_batch_plan_create with
// autotuning on times each on the batch itself and keeps the fastest).
enum BatchShapePolicy {
  kBatchShapeDefault = 0,  // forced shape, else v = 8 and gf_batch_groups(ndst)
  kBatchShapeTable = 1,    // the tuned table's shape for each task's class (single-set tuning)
  kBatchShapeOneGroup = 2, // v = 8, one 8-row group per item
  kBatchShapeSmallV = 3,   // v = 4 for tasks of at most 64 Ki words (more items per small slice)
  kBatchShapePolicies = 4
};
void batch_tile_shape(const par2hip_ctx *ctx, int policy, int ndst, int nsrc, long words, int align_bytes, int *v,
                      int *g);
constexpr int kBatchTpb = 256;  // threads of a batch contraction workgroup
// Many jobs as one batch.  arena: device working memory for the batch (at
// least the sum of batch_job_arena_bytes over the jobs plus
// kBatchArenaExtra), or nullptr for the context's own scratch.
int32_t batch_run(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs, uint8_t *arena, size_t arena_bytes);
size_t batch_job_arena_bytes(const par2hip_job &job);
constexpr size_t kBatchArenaExtra = 512;
// Nibble tables of every task (npairs = all tasks' (8-row group, source)
// pairs, numbered by GfTaskDev::pair0), then the persistent kernel works
// through `items` (sorted by cost) with an atomic cursor.
int32_t launch_tables_batch(par2hip_ctx *ctx, hipStream_t s, const GfTaskDev *d_tasks, int ntasks, long npairs);
int32_t launch_gf_batch(par2hip_ctx *ctx, hipStream_t s, const GfTaskDev *d_tasks, const GfItem *d_items,
                        int nitems, unsigned *d_cursor, const char *tag);
// coefficient matrices of many encode jobs: coeff[r][d] = exp[(log_d * e_r) mod 65535]
struct CoeffJob {
  uint16_t *coeff;
  const uint32_t *src_log, *exps;
  int nrow, nsrc;
};
int32_t launch_coeff_batch(par2hip_ctx *ctx, hipStream_t s, const CoeffJob *d_jobs, int njobs, long max_entries);
// repair jobs: closed-form inverse, V rows and the inverse copy, many at once
struct InvJob {
  uint16_t *aug;          // m x 2m workspace; A^-1 lands in the right half
  const uint32_t *miss_log;
  const uint32_t *plog, *exps;
  uint16_t *vand;         // m rows of np entries, stride vs
  uint16_t *coeff;        // m rows, stride cs: [ fold output | A^-1 ]
  uint64_t vs, cs;
  uint32_t e0;
  int m, np;
};
int32_t launch_inverse_batch(par2hip_ctx *ctx, hipStream_t s, const InvJob *d_jobs, int njobs, int max_m,
                             long max_vand);
int32_t run_gf_lds(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p);
// Variant T with its nibble tables in a caller-provided workspace of at least
// gf_tab_entries(ndst, nsrc) * 16 bytes and the fixed heuristic's (or the
// forced) tile shape: no allocation, no tuning runs.
int32_t run_gf_lds_fixed(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, uint64_t *tabs);
// The main stream forks into the side stream (MD5 chains) and the rest
// stream (ctx->cur), and joins them back (par2hip_api.hip).
int32_t fork_overlap(par2hip_ctx *ctx);
int32_t join_overlap(par2hip_ctx *ctx);
int align_words_of(uintptr_t p);
// Bytes of the per-device workspace of par2hip_encode_dev_noalloc (a
// file-scope device array in noalloc.hip) and its set-up: the array's device
// address on the context's device, looked up once per device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
int32_t noalloc_init(par2hip_ctx *ctx);
int32_t run_gf_mfma(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, bool int8);
// Kernel used when the caller leaves the choice to the library and the call
// has no plan of its own (an encode beside MD5 chains plans kernel and CUs):
// the LDS-table contraction, the faster one alone above 4 outputs (run_gf
// takes the FP4 form's 4-output tile at or below).
constexpr int kAutoGfKernel = 1;

// Slice hashes.
// Largest slice (exclusive) the MD5 kernels take: they count 64-byte blocks
// in 32 bits with room for the prefetch distance.  Every entry point that
// hashes refuses larger slices with PAR2HIP_EINVAL before any launch.
constexpr uint64_t kMd5MaxSliceBytes = uint64_t(1) << 37;
constexpr int kCrcLanes = 256;  // CRC-32: lanes (one workgroup) per slice
// CRC-32 row layout: lane l reads the kCrcPiece bytes at l * kCrcPiece of each
// kCrcRow-byte row, so neighbouring lanes read neighbouring bytes
constexpr uint64_t kCrcPiece = 64, kCrcRow = kCrcLanes * kCrcPiece;
// Per-slice-size constants of the CRC-32 tree combine.
struct CrcParams {
  uint64_t chunk;          // bytes per lane (multiple of 16)
  uint64_t pad;            // leading virtual zero bytes
  uint32_t level[8];       // x^(8*chunk*2^k) mod P, k = 0..7
  uint32_t init_term;      // 0xFFFFFFFF * x^(8*slice_size) mod P
  uint32_t rows;           // whole kCrcRow rows of the slice for the row layout (0: none)
  uint32_t row_level[8];   // x^(8*kCrcPiece*2^k) mod P, k = 0..7
  // the bytes after the rows (fewer than kCrcRow), one chunk per lane:
  uint64_t t_chunk;        // bytes per lane (multiple of 16; 0: no tail)
  uint64_t t_pad;          // leading virtual zero bytes of the tail
  uint32_t t_level[8];     // x^(8*t_chunk*2^k) mod P, k = 0..7
  uint32_t t_shift;        // x^(8*tail bytes) mod P
};
CrcParams crc_params_for(uint64_t slice_size);
// Pointers read from device memory are generic to the compiler, which then
// emits flat memory operations: their completion count is shared with LDS and
// scalar loads, so every wait for a table lookup also waits for HBM.  Typing
// them in the global address space gives global operations with their own
// counter.
#define P2H_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ P2H_GLOBAL T *gptr(T *p) {
  return (P2H_GLOBAL T *)p;
}
// Native vector types (HIP's uint2/uint4 are classes whose operators take
// generic references) for single vector accesses through global pointers.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// One slice of a batched hash pass (device-resident list).
struct HashEntry {
  const uint8_t *data;  // slice bytes
  uint64_t len;         // real bytes (zeros follow up to size)
  uint64_t size;        // padded slice size
  uint32_t *md5;        // 4 words out (nullable for CRC-only lists)
  uint32_t *crc;        // 1 word out
  uint32_t crc_params;  // index into the CrcParams table
  uint32_t pad_;
};
int32_t launch_md5_list(par2hip_ctx *ctx, hipStream_t s, const HashEntry *d_list, size_t n, bool aligned4,
                        bool high_priority);
int32_t launch_crc32_list(par2hip_ctx *ctx, hipStream_t s, const HashEntry *d_list, size_t n,
                          const CrcParams *d_params);
// entries below this size take the lane-per-slice CRC kernel
constexpr uint64_t kCrcLaneMaxBytes = 64 * 1024;
int32_t launch_crc32_list_lane(par2hip_ctx *ctx, hipStream_t s, const HashEntry *d_list, size_t n);

// high_priority: the MD5 chains are the critical path of the caller, so their
// waves issue ahead of co-resident waves of other kernels.
// Whole-set encode on device buffers (par2hip_api.hip), queued on ctx->stream;
// recovery row i at d_rec + i * rec_stride (0: slice_size).
int32_t encode_dev_impl(par2hip_ctx *ctx, const uint8_t *d_slices, size_t n, size_t slice_size,
                        const uint32_t *exps, size_t r, uint8_t *d_rec, uint32_t *d_md5, uint32_t *d_crc,
                        size_t rec_stride);
int32_t launch_md5(par2hip_ctx *ctx, hipStream_t s, const uint8_t *d_data, size_t n,
                   size_t slice_size, size_t stride, const uint64_t *d_lens, uint32_t *d_md5,
                   bool high_priority = false);
int32_t launch_crc32(par2hip_ctx *ctx, hipStream_t s, const uint8_t *d_data, size_t n,
                     size_t slice_size, size_t stride, const uint64_t *d_lens, uint32_t *d_crc);
int32_t crc_tables_init(par2hip_ctx *ctx);

// Gauss-Jordan on an m x 2m workspace [A | I] (device, row-major u16);
// writes 0 or 1 (singular) into *d_status.
int32_t launch_gauss_jordan(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, int m,
                            int *d_status);
// A[i][k] = exp[(miss_log[k]*exps[i]) mod 65535] into the left half of aug, I on the right.
int32_t launch_build_aug(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, const uint32_t *d_miss_log,
                         const uint32_t *d_exps, int m);
// Closed-form inverse for consecutive exponents e0, e0+1, ...: writes A^-1
// into the right half of the m x 2m workspace.
constexpr int kVandMaxM = 1024;  // [tables | P] of 130 m bytes in LDS
int32_t launch_vandermonde_inverse(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, const uint32_t *d_miss_log,
                                   uint32_t e0, int m);
// vand[i][j] = c_{present_j}^{e_i} (row stride vs words).
int32_t launch_build_vand(par2hip_ctx *ctx, hipStream_t s, uint16_t *vand, size_t vs,
                          const uint32_t *d_present_log, const uint32_t *d_exps, int m, int npresent);
// coeff[k][np + i] = inv[k][i] (inv = right half of aug); row stride cs.
int32_t launch_copy_inverse(par2hip_ctx *ctx, hipStream_t s, uint16_t *coeff, size_t cs, const uint16_t *d_aug,
                            int m, int npresent);

}  // namespace par2hip
