// par2hip_api.hip -- the C ABI (include/par2hip.h): context management,
// parameter upload, and the encode / repair / hash / strategy entry points.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>

#include "common.hpp"

namespace par2hip {

HostField::HostField() {
  uint32_t b = 1;
  for (uint32_t i = 0; i < kOrder; i++) {
    exp[i] = (uint16_t)b;
    exp[i + kOrder] = (uint16_t)b;
    log[b] = (uint16_t)i;
    b <<= 1;
    if (b & 0x10000u) b ^= kPoly;
  }
  log[0] = 0;
}

const HostField &field() {
  static const HostField *f = new HostField();
  return *f;
}

std::vector<uint32_t> constant_logs(size_t count) {
  static std::mutex mu;
  static std::vector<uint32_t> logs;
  std::lock_guard<std::mutex> g(mu);
  uint32_t n = logs.empty() ? 1 : logs.back() + 1;
  while (logs.size() < count) {
    if (n % 3 && n % 5 && n % 17 && n % 257) logs.push_back(n);
    n++;
  }
  return std::vector<uint32_t>(logs.begin(), logs.begin() + count);
}

int32_t fail(par2hip_ctx *ctx, int32_t code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int32_t hip_fail(par2hip_ctx *ctx, hipError_t e, const char *what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return fail(ctx, e == hipErrorOutOfMemory ? PAR2HIP_ENOMEM : PAR2HIP_EHIP, buf);
}

int32_t scratch(par2hip_ctx *ctx, int role, size_t bytes, void **out) {
  DevBuf &b = ctx->scratch[role];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // the buffer may still be read by queued work
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->side));
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->rest));
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->wide));
      for (hipStream_t q : ctx->narrow)
        if (q) P2H_CHECK(ctx, hipStreamSynchronize(q));
      P2H_CHECK(ctx, hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)4096);
    want = (want + 4095) & ~(size_t)4095;
    P2H_CHECK(ctx, hipMalloc(&b.ptr, want));
    b.bytes = want;
  }
  *out = b.ptr;
  return PAR2HIP_OK;
}

int32_t host_stage(par2hip_ctx *ctx, size_t bytes, uint8_t **out) {
  if (ctx->host_stage_bytes < bytes) {
    if (ctx->host_stage) {
      // copies from the old buffer may still be queued
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      P2H_CHECK(ctx, hipHostFree(ctx->host_stage));
      ctx->host_stage = nullptr;
      ctx->host_stage_bytes = 0;
    }
    const size_t want = (std::max(bytes, (size_t)1 << 20) + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
    P2H_CHECK(ctx, hipHostMalloc((void **)&ctx->host_stage, want, hipHostMallocDefault));
    ctx->host_stage_bytes = want;
  }
  *out = ctx->host_stage;
  return PAR2HIP_OK;
}

void copy_parallel(const std::vector<CopyJob> &jobs, size_t max_threads) {
  std::vector<CopyJob> pieces;
  const uint64_t kPiece = 8u << 20;
  for (const CopyJob &j : jobs)
    for (uint64_t o = 0; o < j.n; o += kPiece) pieces.push_back({j.dst + o, j.src + o, std::min(kPiece, j.n - o)});
  std::atomic<size_t> next{0};
  auto worker = [&] {
    for (size_t i; (i = next++) < pieces.size();) memcpy(pieces[i].dst, pieces[i].src, pieces[i].n);
  };
  const size_t nt = std::min(pieces.size(), max_threads);
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; t++) th.emplace_back(worker);
  worker();
  for (auto &t : th) t.join();
}

namespace {
constexpr size_t kStageChunk = (size_t)64 << 20;  // bytes per staging half
constexpr size_t kStageDirect = (size_t)4 << 20;  // below this a plain (pageable) copy
constexpr size_t kStageThreads = 8;
}  // namespace

int32_t h2d_staged(par2hip_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
  if (bytes < kStageDirect) {
    if (bytes) P2H_CHECK(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return PAR2HIP_OK;
  }
  const size_t half = std::min(bytes, kStageChunk);
  uint8_t *stage;
  int32_t rc;
  if ((rc = host_stage(ctx, 2 * half, &stage))) return rc;
  const uint8_t *src = static_cast<const uint8_t *>(h_src);
  uint8_t *dst = static_cast<uint8_t *>(d_dst);
  size_t k = 0;
  for (size_t off = 0; off < bytes; off += half, k++) {
    const size_t len = std::min(half, bytes - off);
    uint8_t *buf = stage + (k & 1) * half;
    // the half's last DMA (chunk k-2, or a previous call's) has read it
    P2H_CHECK(ctx, hipEventSynchronize(ctx->stage_ev[k & 1]));
    copy_parallel({{buf, src + off, len}}, kStageThreads);
    P2H_CHECK(ctx, hipMemcpyAsync(dst + off, buf, len, hipMemcpyHostToDevice, ctx->stream));
    P2H_CHECK(ctx, hipEventRecord(ctx->stage_ev[k & 1], ctx->stream));
  }
  return PAR2HIP_OK;
}

int32_t d2h_staged(par2hip_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
  if (bytes < kStageDirect) {
    if (bytes) {
      P2H_CHECK(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return PAR2HIP_OK;
  }
  const size_t half = std::min(bytes, kStageChunk);
  uint8_t *stage;
  int32_t rc;
  if ((rc = host_stage(ctx, 2 * half, &stage))) return rc;
  const uint8_t *src = static_cast<const uint8_t *>(d_src);
  uint8_t *dst = static_cast<uint8_t *>(h_dst);
  const size_t nchunks = (bytes + half - 1) / half;
  auto issue = [&](size_t k) -> int32_t {
    const size_t off = k * half, len = std::min(half, bytes - off);
    P2H_CHECK(ctx, hipMemcpyAsync(stage + (k & 1) * half, src + off, len, hipMemcpyDeviceToHost, ctx->stream));
    P2H_CHECK(ctx, hipEventRecord(ctx->stage_ev[k & 1], ctx->stream));
    return PAR2HIP_OK;
  };
  if ((rc = issue(0))) return rc;
  for (size_t k = 0; k < nchunks; k++) {
    // the half chunk k+1 goes to held chunk k-1, drained in the previous pass
    if (k + 1 < nchunks && (rc = issue(k + 1))) return rc;
    P2H_CHECK(ctx, hipEventSynchronize(ctx->stage_ev[k & 1]));
    const size_t off = k * half, len = std::min(half, bytes - off);
    copy_parallel({{dst + off, stage + (k & 1) * half, len}}, kStageThreads);
  }
  return PAR2HIP_OK;
}

int32_t params_begin(par2hip_ctx *ctx, size_t bytes, uint8_t *target) {
  ParamArena &a = ctx->arena;
  if (a.copied) P2H_CHECK(ctx, hipEventSynchronize(a.copied));
  if (a.cap < bytes) {
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->side));
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->rest));
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->wide));
    for (hipStream_t q : ctx->narrow)
      if (q) P2H_CHECK(ctx, hipStreamSynchronize(q));
    if (a.host) (void)hipHostFree(a.host);
    if (a.dev) (void)hipFree(a.dev);
    a.host = nullptr;
    a.dev = nullptr;
    size_t cap = std::max(bytes, (size_t)1 << 20);
    P2H_CHECK(ctx, hipHostMalloc((void **)&a.host, cap, hipHostMallocDefault));
    P2H_CHECK(ctx, hipMalloc((void **)&a.dev, cap));
    a.cap = cap;
  }
  a.used = 0;
  a.target = target ? target : a.dev;
  return PAR2HIP_OK;
}

void *params_alloc(par2hip_ctx *ctx, size_t bytes, void **dev) {
  ParamArena &a = ctx->arena;
  size_t off = (a.used + 255) & ~(size_t)255;
  a.used = off + bytes;
  *dev = a.target + off;
  return a.host + off;
}

int32_t params_commit(par2hip_ctx *ctx) {
  ParamArena &a = ctx->arena;
  if (a.used) {
    P2H_CHECK(ctx, hipMemcpyAsync(a.target, a.host, a.used, hipMemcpyHostToDevice, ctx->cur));
    P2H_CHECK(ctx, hipEventRecord(a.copied, ctx->cur));
  }
  return PAR2HIP_OK;
}

static size_t params_size(std::initializer_list<size_t> parts) {
  size_t t = 0;
  for (size_t p : parts) t += (p + 255) & ~(size_t)255;
  return t + 256;
}

static hipEvent_t pool_event(par2hip_ctx *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int timing_start(par2hip_ctx *ctx, const char *name, hipStream_t s, hipEvent_t *stop) {
  *stop = nullptr;
  if (!ctx->profile) return -1;
  int slot = -1;
  for (size_t i = 0; i < ctx->stats.size(); i++)
    if (ctx->stats[i].name == name) slot = (int)i;
  if (slot < 0) {
    ctx->stats.push_back(KernelStat{name, 0, 0});
    slot = (int)ctx->stats.size() - 1;
  }
  hipEvent_t start = pool_event(ctx);
  *stop = pool_event(ctx);
  if (!start || !*stop) return -1;
  (void)hipEventRecord(start, s);
  ctx->pending.push_back(PendingTiming{slot, start, *stop});
  return slot;
}

void timing_stop(par2hip_ctx *ctx, int slot, hipStream_t s, hipEvent_t stop) {
  if (slot < 0 || !stop) return;
  (void)hipEventRecord(stop, s);
}

}  // namespace par2hip

using namespace par2hip;

namespace {

struct Guard {
  par2hip_ctx *ctx;
  std::lock_guard<std::recursive_mutex> lk;
  explicit Guard(par2hip_ctx *c) : ctx(c), lk(c->mu) { (void)hipSetDevice(c->device); }
};

int g_ndev = -1;
std::once_flag g_dev_once;

void probe_devices() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  int ok = 0;
  for (int i = 0; i < n; i++) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, i) == hipSuccess && strncmp(p.gcnArchName, "gfx950", 6) == 0) ok++;
  }
  g_ndev = ok;
}

// Overlapped section: the side stream (MD5 chains, a few reserved CUs) and
// the rest stream (all other CUs) both start after everything queued on the
// main stream; the call's main work then goes to `rest` until join.
// work: the stream for the call's main work (default: rest, every CU outside
// the side stream's mask).
int32_t fork_side(par2hip_ctx *ctx, hipStream_t work = nullptr) {
  if (!work) work = ctx->rest;
  P2H_CHECK(ctx, hipEventRecord(ctx->fork, ctx->stream));
  P2H_CHECK(ctx, hipStreamWaitEvent(ctx->side, ctx->fork, 0));
  if (work != ctx->stream) P2H_CHECK(ctx, hipStreamWaitEvent(work, ctx->fork, 0));
  ctx->cur = ctx->rest_cur = work;
  return PAR2HIP_OK;
}

int32_t join_side(par2hip_ctx *ctx) {
  P2H_CHECK(ctx, hipEventRecord(ctx->join, ctx->side));
  P2H_CHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->join, 0));
  hipStream_t work = ctx->rest_cur ? ctx->rest_cur : ctx->rest;
  if (work != ctx->stream) {
    P2H_CHECK(ctx, hipEventRecord(ctx->join2, work));
    P2H_CHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->join2, 0));
  }
  ctx->cur = ctx->stream;
  ctx->rest_cur = nullptr;
  return PAR2HIP_OK;
}

// Forks on construction, joins on destruction (also on early error returns).
struct Overlap {
  par2hip_ctx *ctx;
  int32_t rc;
  explicit Overlap(par2hip_ctx *c, hipStream_t work = nullptr) : ctx(c) { rc = fork_side(c, work); }
  ~Overlap() { (void)join_side(ctx); }
};

// Kernel and CUs of an encode contraction that overlaps the MD5 chains of
// the same call on the side CUs.  Alone, the LDS-table contraction is the
// faster one; beside the chains the chip is power-bound: the LDS form holds
// every CU busy on lookups and pulls the shader clock down, and the chains,
// the step's critical path, slow with it.  The bit-sliced FP4 form on only as
// many CUs as it needs to finish under the chains (here: the narrowest of the
// narrow streams and the full complement that brings its estimate under
// kPlanSlack of the chains' estimated time) leaves them the clock.  1 GiB /
// R=102 with 32 MD5 CUs and the helper-wave chains (8.2 ms alone): the FP4
// form on the 224 other CUs 8.39-8.48 ms per step (chains 8.28-8.37 ms,
// contraction 6.6 ms, about 1.15x its estimate), the LDS form on the same
// CUs 8.91-8.96 ms (chains 8.80-8.86 ms), two alternating rounds on one box.
// With the plain chains (9.4 ms) the FP4 form on 128 CUs took 10.88 ms per
// step (its contraction outlasted the chains), on 160 or 192 CUs 9.80 ms.  If
// no width fits, the contraction is the critical path and the LDS form on
// every other CU runs.
constexpr double kPlanSlack = 0.85;
struct GfPlan {
  int kernel = 0;             // 0: the context's choice
  hipStream_t work = nullptr;  // nullptr: rest
};
static GfPlan plan_overlapped_gf(const par2hip_ctx *ctx, const void *d_slices, size_t n, size_t slice_size, size_t r,
                                 int align_words) {
  GfPlan g;
  if (!ctx->masked || ctx->gf_kernel != 0 || r == 0) return g;
  g.kernel = 1;
  if (align_words < 8) return g;
  const double md5_ms = md5_chains_ms(n, slice_size, 4 * ctx->hash_cus,
                                     md5_helper_form(ctx, ctx->side, d_slices, n, slice_size, slice_size, false));
  const double macs = 256.0 * (double)r * (double)n * (double)(slice_size / 2);
  for (int i = 0; i <= 3; i++) {
    const int cus = i < 3 ? ctx->narrow_cus[i] : ctx->num_cus - ctx->hash_cus;
    if (cus <= 0 || (i < 3 && !ctx->narrow[i])) continue;
    if (macs / (kFp4MacPerCuSec * cus) * 1e3 <= kPlanSlack * md5_ms) {
      g.kernel = 2;
      g.work = i < 3 ? ctx->narrow[i] : nullptr;
      break;
    }
  }
  return g;
}

}  // namespace

int par2hip::align_words_of(uintptr_t p) {
  int a = 8;
  while (a > 1 && (p % (2 * (uintptr_t)a)) != 0) a >>= 1;
  return a;
}

namespace {

// Common encode body: device buffers in, queued on ctx->stream.  Recovery
// row i goes to d_rec + i * rec_stride (rec_stride 0: slice_size).
int32_t encode_body(par2hip_ctx *ctx, const uint8_t *d_slices, size_t n, size_t slice_size,
                    const uint32_t *exps, size_t r, uint8_t *d_rec, uint32_t *d_md5,
                    uint32_t *d_crc, size_t rec_stride) {
  if (rec_stride == 0) rec_stride = slice_size;
  if (n == 0 || slice_size == 0 || (slice_size & 1))
    return fail(ctx, PAR2HIP_EINVAL, "need n >= 1 and an even, non-zero slice size");
  if (n > 32768) return fail(ctx, PAR2HIP_EINVAL, "more than 32768 input slices repeat PAR2 constants");
  if (r > 65535 || (r && !exps)) return fail(ctx, PAR2HIP_EINVAL, "bad recovery exponent list");
  if (d_md5 && slice_size >= kMd5MaxSliceBytes) return fail(ctx, PAR2HIP_EINVAL, "hashed slice of 128 GiB or more");
  const size_t words = slice_size / 2;
  const std::vector<uint32_t> logs = constant_logs(n);
  void *coeff = nullptr;
  int32_t rc;
  if (r && (rc = scratch(ctx, kScratchCoeff, r * n * 2, &coeff))) return rc;
  if ((rc = params_begin(ctx, params_size({n * 4, r * 4, n * 8, r * 8})))) return rc;
  void *d_logs, *d_exps, *d_srcs, *d_dsts;
  uint32_t *h_logs = (uint32_t *)params_alloc(ctx, n * 4, &d_logs);
  uint32_t *h_exps = (uint32_t *)params_alloc(ctx, std::max<size_t>(r, 1) * 4, &d_exps);
  const uint16_t **h_srcs = (const uint16_t **)params_alloc(ctx, n * 8, &d_srcs);
  uint16_t **h_dsts = (uint16_t **)params_alloc(ctx, std::max<size_t>(r, 1) * 8, &d_dsts);
  memcpy(h_logs, logs.data(), n * 4);
  for (size_t i = 0; i < r; i++) h_exps[i] = exps[i] % kOrder;  // c^e = c^(e mod 65535)
  for (size_t d = 0; d < n; d++) h_srcs[d] = (const uint16_t *)(d_slices + d * slice_size);
  for (size_t i = 0; i < r; i++) h_dsts[i] = (uint16_t *)(d_rec + i * rec_stride);
  if ((rc = params_commit(ctx))) return rc;

  // MD5 is a long serial chain per slice: it runs on the side stream's CUs
  // from the start while CRC-32 and the contraction use the other CUs.
  const int align = std::min({align_words_of((uintptr_t)d_slices), align_words_of((uintptr_t)d_rec),
                              align_words_of(slice_size), align_words_of(rec_stride)});
  const GfPlan plan = d_md5 ? plan_overlapped_gf(ctx, d_slices, n, slice_size, r, align) : GfPlan{};
  std::unique_ptr<Overlap> ov;
  if (d_md5) {
    ov.reset(new Overlap(ctx, plan.work));
    if (ov->rc) return ov->rc;
    if ((rc = launch_md5(ctx, ctx->side, d_slices, n, slice_size, slice_size, nullptr, d_md5, !ctx->masked)))
      return rc;
  }
  if (d_crc && (rc = launch_crc32(ctx, ctx->cur, d_slices, n, slice_size, slice_size, nullptr, d_crc)))
    return rc;
  if (r) {
    if ((rc = launch_coeff_from_logs(ctx, ctx->cur, (uint16_t *)coeff, (const uint32_t *)d_logs,
                                     (const uint32_t *)d_exps, (int)r, (int)n)))
      return rc;
    GfProblem p;
    p.srcs = (const uint16_t *const *)d_srcs;
    p.dsts = (uint16_t *const *)d_dsts;
    p.coeff = (const uint16_t *)coeff;
    p.nsrc = (int)n;
    p.ndst = (int)r;
    p.words = words;
    p.accumulate = false;
    p.align_words = align;
    p.kernel = plan.kernel;
    p.beside_chains = d_md5 != nullptr;
    if ((rc = run_gf(ctx, ctx->cur, p))) return rc;
  }
  return PAR2HIP_OK;
}

// Repair body on device buffers.  Sources of the contraction are the present
// data slices followed by the m recovery slices.
// rec_rows (nullable): which rows of d_recovery hold the m recovery slices
// with exponents rec_exps[0..m) (default: rows 0..m-1).
int32_t reconstruct_dev_impl(par2hip_ctx *ctx, uint8_t *d_slices, size_t n, size_t slice_size,
                             const uint32_t *missing, size_t m, const uint8_t *d_recovery,
                             const uint32_t *rec_exps, uint8_t *d_out, bool in_place,
                             const uint32_t *rec_rows = nullptr) {
  if (n == 0 || slice_size == 0 || (slice_size & 1))
    return fail(ctx, PAR2HIP_EINVAL, "need n >= 1 and an even, non-zero slice size");
  if (n > 32768) return fail(ctx, PAR2HIP_EINVAL, "more than 32768 input slices repeat PAR2 constants");
  if (m == 0) return PAR2HIP_OK;
  if (m > n) return fail(ctx, PAR2HIP_EINVAL, "more erasures than slices");
  std::vector<char> is_missing(n, 0);
  for (size_t k = 0; k < m; k++) {
    if (missing[k] >= n) return fail(ctx, PAR2HIP_EINVAL, "erasure index out of range");
    if (is_missing[missing[k]]) return fail(ctx, PAR2HIP_EINVAL, "duplicate erasure index");
    is_missing[missing[k]] = 1;
  }
  for (size_t i = 0; i < m; i++)
    for (size_t j = i + 1; j < m; j++)
      if (rec_exps[i] == rec_exps[j]) return fail(ctx, PAR2HIP_ESINGULAR, "Singular matrix (repeated recovery exponent)");
  const size_t np = n - m;
  const size_t words = slice_size / 2;
  const std::vector<uint32_t> logs = constant_logs(n);
  // device workspaces first: the one parameter block of this call holds
  // pointers into them.  C = [A^-1 V | A^-1] over [present | recovery], rows
  // padded to a multiple of 8 entries; V rows padded likewise.
  const size_t cs = (np + m + 7) & ~(size_t)7, vs = (std::max<size_t>(np, 1) + 7) & ~(size_t)7;
  void *aug, *coeff, *vand;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchGj, m * 2 * m * 2, &aug))) return rc;
  if ((rc = scratch(ctx, kScratchCoeff, m * cs * 2, &coeff))) return rc;
  if ((rc = scratch(ctx, kScratchVand, m * vs * 2, &vand))) return rc;
  if ((rc = params_begin(ctx, params_size({m * 4, m * 4, np * 4, (np + m) * 8, m * 8, m * 8, m * 8})))) return rc;
  void *d_mlog, *d_exps, *d_plog, *d_srcs, *d_dsts, *d_fsrc, *d_fdst;
  uint32_t *h_mlog = (uint32_t *)params_alloc(ctx, m * 4, &d_mlog);
  uint32_t *h_exps = (uint32_t *)params_alloc(ctx, m * 4, &d_exps);
  uint32_t *h_plog = (uint32_t *)params_alloc(ctx, std::max<size_t>(np, 1) * 4, &d_plog);
  const uint16_t **h_srcs = (const uint16_t **)params_alloc(ctx, (np + m) * 8, &d_srcs);
  uint16_t **h_dsts = (uint16_t **)params_alloc(ctx, m * 8, &d_dsts);
  const uint16_t **h_fsrc = (const uint16_t **)params_alloc(ctx, m * 8, &d_fsrc);
  uint16_t **h_fdst = (uint16_t **)params_alloc(ctx, m * 8, &d_fdst);
  for (size_t k = 0; k < m; k++) h_mlog[k] = logs[missing[k]];
  for (size_t i = 0; i < m; i++) h_exps[i] = rec_exps[i] % kOrder;
  size_t j = 0;
  for (size_t d = 0; d < n; d++)
    if (!is_missing[d]) {
      h_plog[j] = logs[d];
      h_srcs[j] = (const uint16_t *)(d_slices + d * slice_size);
      j++;
    }
  for (size_t i = 0; i < m; i++)
    h_srcs[np + i] = (const uint16_t *)(d_recovery + (size_t)(rec_rows ? rec_rows[i] : i) * slice_size);
  for (size_t k = 0; k < m; k++)
    h_dsts[k] = (uint16_t *)(in_place ? d_slices + (size_t)missing[k] * slice_size : d_out + k * slice_size);
  for (size_t i = 0; i < m; i++) {
    h_fsrc[i] = (const uint16_t *)vand + i * vs;
    h_fdst[i] = (uint16_t *)coeff + i * cs;
  }
  if ((rc = params_commit(ctx))) return rc;

  if ((rc = launch_build_vand(ctx, ctx->cur, (uint16_t *)vand, vs, (const uint32_t *)d_plog,
                              (const uint32_t *)d_exps, (int)m, (int)np)))
    return rc;
  // consecutive exponents (the lowest ones, as the reference picks them when
  // no volume is missing) have a closed-form inverse; otherwise Gauss-Jordan
  bool consecutive = m <= (size_t)kVandMaxM;
  for (size_t i = 1; i < m && consecutive; i++) consecutive = rec_exps[i] == rec_exps[0] + i;
  if (consecutive && rec_exps[0] + m - 1 >= kOrder) consecutive = false;
  if (consecutive) {
    if ((rc = launch_vandermonde_inverse(ctx, ctx->cur, (uint16_t *)aug, (const uint32_t *)d_mlog, rec_exps[0],
                                         (int)m)))
      return rc;
  } else {
    if ((rc = launch_build_aug(ctx, ctx->cur, (uint16_t *)aug, (const uint32_t *)d_mlog,
                               (const uint32_t *)d_exps, (int)m)))
      return rc;
    if ((rc = launch_gauss_jordan(ctx, ctx->cur, (uint16_t *)aug, (int)m, ctx->d_status))) return rc;
    int status = 0;
    P2H_CHECK(ctx, hipMemcpyAsync(&status, ctx->d_status, sizeof status, hipMemcpyDeviceToHost, ctx->cur));
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->cur));
    if (status) return fail(ctx, PAR2HIP_ESINGULAR, "Singular matrix");
  }
  if ((rc = launch_copy_inverse(ctx, ctx->cur, (uint16_t *)coeff, cs, (const uint16_t *)aug, (int)m, (int)np)))
    return rc;
  if (np) {
    // fold: row k of A^-1 V = XOR_i inv[k][i] * V_i -- the contraction itself,
    // with V's rows as m sources of np words and C's rows as outputs
    GfProblem f;
    f.srcs = (const uint16_t *const *)d_fsrc;
    f.dsts = (uint16_t *const *)d_fdst;
    f.coeff = (const uint16_t *)aug + m;  // A^-1, row stride 2m
    f.ldc = (int)(2 * m);
    f.nsrc = (int)m;
    f.ndst = (int)m;
    f.words = np;
    f.accumulate = false;
    f.align_words = 8;
    f.tag = "gf_fold";
    if ((rc = run_gf(ctx, ctx->cur, f))) return rc;
  }
  GfProblem p;
  p.srcs = (const uint16_t *const *)d_srcs;
  p.dsts = (uint16_t *const *)d_dsts;
  p.coeff = (const uint16_t *)coeff;
  p.ldc = (int)cs;
  p.nsrc = (int)(np + m);
  p.ndst = (int)m;
  p.words = words;
  p.accumulate = false;
  p.align_words = std::min({align_words_of((uintptr_t)d_slices), align_words_of((uintptr_t)d_recovery),
                            in_place ? 8 : align_words_of((uintptr_t)d_out), align_words_of(slice_size)});
  // sources never include the missing rows, so writing them in place is safe
  return run_gf(ctx, ctx->cur, p);
}

int32_t repair_set_impl(par2hip_ctx *ctx, uint8_t *d_slices, size_t n, size_t slice_size,
                        const uint8_t *expected_md5, const uint32_t *expected_crc, const uint8_t *d_recovery,
                        const uint32_t *rec_exps, size_t nrec, uint8_t *d_out, uint32_t *damaged_out,
                        size_t *ndamaged) {
  if (n == 0 || slice_size == 0 || (slice_size & 1))
    return fail(ctx, PAR2HIP_EINVAL, "need n >= 1 and an even, non-zero slice size");
  if (slice_size >= kMd5MaxSliceBytes) return fail(ctx, PAR2HIP_EINVAL, "hashed slice of 128 GiB or more");
  int32_t rc;
  void *dh;
  if ((rc = scratch(ctx, kScratchHashD, n * 20 + 16, &dh))) return rc;
  uint32_t *d_md5 = (uint32_t *)dh, *d_crc = (uint32_t *)dh + 4 * n;
  // MD5 (long serial chains) on the side stream's CUs, CRC-32 and the
  // rebuild on the other CUs
  Overlap ov(ctx);
  if (ov.rc) return ov.rc;
  if ((rc = launch_md5(ctx, ctx->side, d_slices, n, slice_size, slice_size, nullptr, d_md5))) return rc;
  if ((rc = launch_crc32(ctx, ctx->cur, d_slices, n, slice_size, slice_size, nullptr, d_crc))) return rc;
  std::vector<uint32_t> crc(n);
  P2H_CHECK(ctx, hipMemcpyAsync(crc.data(), d_crc, n * 4, hipMemcpyDeviceToHost, ctx->cur));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->cur));
  std::vector<char> bad(n, 0);
  std::vector<uint32_t> dmg;
  for (size_t i = 0; i < n; i++)
    if (crc[i] != expected_crc[i]) {
      bad[i] = 1;
      dmg.push_back((uint32_t)i);
    }
  // recovery slices in ascending exponent order (repair.rs:584-604)
  std::vector<uint32_t> order(nrec);
  for (size_t i = 0; i < nrec; i++) order[i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rec_exps[a] < rec_exps[b]; });
  auto rebuild = [&](const std::vector<uint32_t> &d) -> int32_t {
    if (d.empty()) return PAR2HIP_OK;
    const size_t m = d.size();
    std::vector<uint32_t> rows(order.begin(), order.begin() + m), exps(m);
    for (size_t k = 0; k < m; k++) exps[k] = rec_exps[rows[k]];
    return reconstruct_dev_impl(ctx, d_slices, n, slice_size, d.data(), m, d_recovery, exps.data(), d_out,
                                d_out == nullptr, rows.data());
  };
  bool rebuilt = false;
  if (dmg.size() <= nrec) {
    // speculative: MD5 may still add slices.  An in-place rebuild rewrites the
    // CRC-damaged rows while the MD5 pass still reads them; their digests are
    // never used (only slices whose CRC matched are confirmed by MD5).
    if ((rc = rebuild(dmg))) return rc;
    rebuilt = true;
  }
  // MD5 confirmation of every slice whose CRC matched (verify.rs:631-645)
  std::vector<uint8_t> md5(n * 16);
  P2H_CHECK(ctx, hipMemcpyAsync(md5.data(), d_md5, n * 16, hipMemcpyDeviceToHost, ctx->side));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->side));
  bool more = false;
  for (size_t i = 0; i < n; i++)
    if (!bad[i] && memcmp(&md5[i * 16], expected_md5 + i * 16, 16) != 0) {
      bad[i] = 1;
      more = true;
    }
  if (more) {
    dmg.clear();
    for (size_t i = 0; i < n; i++)
      if (bad[i]) dmg.push_back((uint32_t)i);
  }
  for (size_t k = 0; k < dmg.size(); k++) damaged_out[k] = dmg[k];
  *ndamaged = dmg.size();
  if (dmg.size() > nrec) {
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->cur));
    char buf[128];
    snprintf(buf, sizeof buf, "Need %zu recovery blocks, have %zu", dmg.size(), nrec);
    return fail(ctx, PAR2HIP_ERECOVERY, buf);
  }
  if (more || !rebuilt) {
    // the first rebuild (if any) treated a damaged slice as intact: redo it
    if ((rc = rebuild(dmg))) return rc;
  }
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->cur));
  return PAR2HIP_OK;
}

}  // namespace

namespace par2hip {
// entry points of the single-set path used by the batch planner (batch.hip)
int32_t reconstruct_single(par2hip_ctx *ctx, uint8_t *d_slices, size_t n, size_t slice_size, const uint32_t *missing,
                           size_t m, const uint8_t *d_recovery, const uint32_t *rec_exps, uint8_t *d_out) {
  return reconstruct_dev_impl(ctx, d_slices, n, slice_size, missing, m, d_recovery, rec_exps, d_out, false);
}
int32_t encode_dev_impl(par2hip_ctx *ctx, const uint8_t *d_slices, size_t n, size_t slice_size,
                        const uint32_t *exps, size_t r, uint8_t *d_rec, uint32_t *d_md5, uint32_t *d_crc,
                        size_t rec_stride) {
  return encode_body(ctx, d_slices, n, slice_size, exps, r, d_rec, d_md5, d_crc, rec_stride);
}
int32_t fork_overlap(par2hip_ctx *ctx) { return fork_side(ctx); }
int32_t join_overlap(par2hip_ctx *ctx) { return join_side(ctx); }
}  // namespace par2hip

extern "C" {

int32_t par2hip_available(void) {
  std::call_once(g_dev_once, probe_devices);
  return g_ndev;
}

const char *par2hip_strategy_name(void) { return "HIP-gfx950"; }
int32_t par2hip_strategy_priority(void) { return PAR2HIP_PRIORITY_OPTIMAL; }

int32_t par2hip_ctx_create(int32_t device, par2hip_ctx **out) { return par2hip::ctx_create(device, true, out); }

}  // extern "C"

// narrow_streams: also create the narrower CU-masked streams the single-set
// encode plans its contraction on.  Each CU-masked stream holds a hardware
// queue of its own; contexts that only run batches (the server's lanes) skip
// them so that several lanes do not oversubscribe the device's queues.
int32_t par2hip::ctx_create(int32_t device, bool narrow_streams, par2hip_ctx **out) {
  if (!out) return PAR2HIP_EINVAL;
  *out = nullptr;
  if (par2hip_available() <= 0) return PAR2HIP_ENODEV;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PAR2HIP_ENODEV;
  std::unique_ptr<par2hip_ctx> c(new par2hip_ctx());
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return PAR2HIP_ENODEV;
  par2hip_ctx *ctx = c.get();
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cus = prop.multiProcessorCount;
  P2H_CHECK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  P2H_CHECK(ctx, hipStreamCreateWithFlags(&ctx->wide, hipStreamNonBlocking));
  ctx->cur = ctx->stream;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) strncpy(ctx->arch, prop.gcnArchName, sizeof ctx->arch - 1);
  {
    // MD5 chains get CUs of their own; overlapped work uses the complementary
    // mask.  32 CUs hold one 64-chain wave each up to 2048 slices: a chain
    // wave alone on its CU, with that CU's vector cache and memory path to
    // itself, keeps its prefetched lines and runs ~10% faster beside a
    // contraction than two or four waves sharing a CU, and the contraction
    // loses less to that interference than it gains from the extra CUs
    // (1 GiB encode 10.13 -> 9.74 ms, 2000 x 512 KiB repair 5.81 -> 5.10 ms
    // against 8 CUs).  Without CU masks the side stream is an ordinary stream
    // and `rest` is the main stream.
    const int ncu = ctx->num_cus, hash_cus = ncu >= 128 ? 32 : ncu >= 64 ? 8 : 0;
    std::vector<uint32_t> side_mask((ncu + 31) / 32, 0u), rest_mask((ncu + 31) / 32, 0u);
    for (int c = 0; c < ncu; c++) (c < hash_cus ? side_mask : rest_mask)[c / 32] |= 1u << (c % 32);
    bool ok = hash_cus > 0 &&
              hipExtStreamCreateWithCUMask(&ctx->side, (uint32_t)side_mask.size(), side_mask.data()) == hipSuccess;
    if (ok && hipExtStreamCreateWithCUMask(&ctx->rest, (uint32_t)rest_mask.size(), rest_mask.data()) != hipSuccess) {
      (void)hipStreamDestroy(ctx->side);
      ok = false;
    }
    ctx->masked = ok;
    ctx->hash_cus = ok ? hash_cus : 0;
    for (int i = 0; ok && narrow_streams && i < 3; i++) {
      const int k = 128 + 32 * i;  // 128, 160, 192 CUs after the side stream's
      if (hash_cus + k >= ncu) break;
      std::vector<uint32_t> m((ncu + 31) / 32, 0u);
      for (int c = hash_cus; c < hash_cus + k; c++) m[c / 32] |= 1u << (c % 32);
      if (hipExtStreamCreateWithCUMask(&ctx->narrow[i], (uint32_t)m.size(), m.data()) == hipSuccess)
        ctx->narrow_cus[i] = k;
      else
        ctx->narrow[i] = nullptr;
    }
    if (!ok) {
      P2H_CHECK(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
      ctx->rest = ctx->stream;
    }
  }
  P2H_CHECK(ctx, hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
  P2H_CHECK(ctx, hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
  P2H_CHECK(ctx, hipEventCreateWithFlags(&ctx->join2, hipEventDisableTiming));
  for (hipEvent_t &e : ctx->stage_ev) P2H_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  P2H_CHECK(ctx, hipEventCreateWithFlags(&ctx->arena.copied, hipEventDisableTiming));
  const HostField &f = field();
  P2H_CHECK(ctx, hipMalloc(&ctx->d_exp, sizeof f.exp));
  P2H_CHECK(ctx, hipMalloc(&ctx->d_log, sizeof f.log));
  P2H_CHECK(ctx, hipMemcpy(ctx->d_exp, f.exp, sizeof f.exp, hipMemcpyHostToDevice));
  P2H_CHECK(ctx, hipMemcpy(ctx->d_log, f.log, sizeof f.log, hipMemcpyHostToDevice));
  P2H_CHECK(ctx, hipMalloc(&ctx->d_status, 256));
  int32_t rc = crc_tables_init(ctx);
  if (rc) return rc;
  if ((rc = noalloc_init(ctx))) return rc;
  *out = c.release();
  return PAR2HIP_OK;
}

extern "C" {

int32_t par2hip_ctx_destroy(par2hip_ctx *ctx) {
  if (!ctx) return PAR2HIP_EINVAL;
  {
    Guard g(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->side);
    if (ctx->rest != ctx->stream) (void)hipStreamSynchronize(ctx->rest);
    (void)hipStreamSynchronize(ctx->wide);
    for (hipStream_t q : ctx->narrow)
      if (q) (void)hipStreamSynchronize(q);
    for (auto &b : ctx->scratch)
      if (b.ptr) (void)hipFree(b.ptr);
    if (ctx->arena.host) (void)hipHostFree(ctx->arena.host);
    if (ctx->arena.dev) (void)hipFree(ctx->arena.dev);
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    if (ctx->d_exp) (void)hipFree(ctx->d_exp);
    if (ctx->d_log) (void)hipFree(ctx->d_log);
    if (ctx->d_crc_tab) (void)hipFree(ctx->d_crc_tab);
    if (ctx->d_status) (void)hipFree(ctx->d_status);
    for (auto &p : ctx->pending) {
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->fork);
    (void)hipEventDestroy(ctx->join);
    (void)hipEventDestroy(ctx->join2);
    for (hipEvent_t e : ctx->stage_ev)
      if (e) (void)hipEventDestroy(e);
    if (ctx->rest != ctx->stream) (void)hipStreamDestroy(ctx->rest);
    (void)hipStreamDestroy(ctx->wide);
    for (hipStream_t q : ctx->narrow)
      if (q) (void)hipStreamDestroy(q);
    (void)hipEventDestroy(ctx->arena.copied);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->side);
  }
  delete ctx;
  return PAR2HIP_OK;
}

const char *par2hip_last_error(par2hip_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int32_t par2hip_get_device_info(par2hip_ctx *ctx, par2hip_device_info *out) {
  if (!ctx || !out) return PAR2HIP_EINVAL;
  memset(out, 0, sizeof *out);
  out->compute_units = ctx->num_cus;
  out->hash_cus = ctx->hash_cus;
  strncpy(out->arch, ctx->arch, sizeof out->arch - 1);
  return PAR2HIP_OK;
}

int32_t par2hip_dev_alloc(par2hip_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return PAR2HIP_EINVAL;
  Guard g(ctx);
  *dptr = nullptr;
  P2H_CHECK(ctx, hipMalloc(dptr, bytes ? bytes : 1));
  return PAR2HIP_OK;
}

int32_t par2hip_dev_free(par2hip_ctx *ctx, void *dptr) {
  if (!ctx) return PAR2HIP_EINVAL;
  Guard g(ctx);
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  P2H_CHECK(ctx, hipFree(dptr));
  return PAR2HIP_OK;
}

int32_t par2hip_h2d(par2hip_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return PAR2HIP_EINVAL;
  Guard g(ctx);
  int32_t rc = h2d_staged(ctx, dst, src, bytes);
  if (rc) return rc;
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PAR2HIP_OK;
}

int32_t par2hip_d2h(par2hip_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return PAR2HIP_EINVAL;
  Guard g(ctx);
  return d2h_staged(ctx, dst, src, bytes);
}

int32_t par2hip_d2d(par2hip_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return PAR2HIP_EINVAL;
  Guard g(ctx);
  P2H_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return PAR2HIP_OK;
}

int32_t par2hip_sync(par2hip_ctx *ctx) {
  if (!ctx) return PAR2HIP_EINVAL;
  Guard g(ctx);
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->side));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->rest));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->wide));
  for (hipStream_t q : ctx->narrow)
    if (q) P2H_CHECK(ctx, hipStreamSynchronize(q));
  return PAR2HIP_OK;
}

// ---------------------------------------------------- strategy-shaped --
int32_t par2hip_gf_muladd_region(par2hip_ctx *ctx, uint16_t *const *dsts, size_t ndst,
                                 const uint16_t *const *srcs, size_t nsrc,
                                 const uint16_t *const *coeff_rows, const size_t *coeff_lens,
                                 size_t region_offset, size_t region_size) {
  if (!ctx) return PAR2HIP_EINVAL;
  if (ndst == 0 || region_size == 0) return PAR2HIP_OK;
  if (!dsts || (nsrc && (!srcs || !coeff_rows))) return PAR2HIP_EINVAL;
  if (nsrc == 0) return PAR2HIP_OK;
  if (ndst > 65535 || nsrc > 65535) return fail(ctx, PAR2HIP_EINVAL, "too many rows");
  Guard g(ctx);
  const size_t rb = region_size * 2;
  const size_t pitch = (rb + 255) & ~(size_t)255;
  void *din, *dout;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchIn, nsrc * pitch, &din))) return rc;
  if ((rc = scratch(ctx, kScratchOut, ndst * pitch, &dout))) return rc;
  if ((rc = params_begin(ctx, params_size({ndst * nsrc * 2, nsrc * 8, ndst * 8})))) return rc;
  void *d_coeff, *d_srcs, *d_dsts;
  uint16_t *h_coeff = (uint16_t *)params_alloc(ctx, ndst * nsrc * 2, &d_coeff);
  const uint16_t **h_srcs = (const uint16_t **)params_alloc(ctx, nsrc * 8, &d_srcs);
  uint16_t **h_dsts = (uint16_t **)params_alloc(ctx, ndst * 8, &d_dsts);
  // row i applies its first min(len_i, nsrc) entries; sources past a short
  // row's end are not applied (scalar.rs:100, `.enumerate().take(num_srcs)`)
  for (size_t i = 0; i < ndst; i++) {
    const size_t k = coeff_lens ? std::min(coeff_lens[i], nsrc) : nsrc;
    if (k) memcpy(h_coeff + i * nsrc, coeff_rows[i], k * 2);
    memset(h_coeff + i * nsrc + k, 0, (nsrc - k) * 2);
  }
  for (size_t j = 0; j < nsrc; j++) h_srcs[j] = (const uint16_t *)((uint8_t *)din + j * pitch);
  for (size_t i = 0; i < ndst; i++) h_dsts[i] = (uint16_t *)((uint8_t *)dout + i * pitch);
  if ((rc = params_commit(ctx))) return rc;
  for (size_t j = 0; j < nsrc; j++)
    P2H_CHECK(ctx, hipMemcpyAsync((uint8_t *)din + j * pitch, srcs[j] + region_offset, rb,
                                  hipMemcpyHostToDevice, ctx->stream));
  for (size_t i = 0; i < ndst; i++)
    P2H_CHECK(ctx, hipMemcpyAsync((uint8_t *)dout + i * pitch, dsts[i] + region_offset, rb,
                                  hipMemcpyHostToDevice, ctx->stream));
  GfProblem p;
  p.srcs = (const uint16_t *const *)d_srcs;
  p.dsts = (uint16_t *const *)d_dsts;
  p.coeff = (const uint16_t *)d_coeff;
  p.nsrc = (int)nsrc;
  p.ndst = (int)ndst;
  p.words = region_size;
  p.accumulate = true;
  p.align_words = 8;
  if ((rc = run_gf(ctx, ctx->stream, p))) return rc;
  for (size_t i = 0; i < ndst; i++)
    P2H_CHECK(ctx, hipMemcpyAsync(dsts[i] + region_offset, (uint8_t *)dout + i * pitch, rb,
                                  hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PAR2HIP_OK;
}

int32_t par2hip_gf_muladd(par2hip_ctx *ctx, uint16_t *dst, const uint16_t *src, size_t len,
                          uint16_t scalar) {
  if (!ctx) return PAR2HIP_EINVAL;
  if (len == 0 || scalar == 0) return PAR2HIP_OK;
  if (!dst || !src) return PAR2HIP_EINVAL;
  const uint16_t *cr[1] = {&scalar};
  uint16_t *d[1] = {dst};
  const uint16_t *s[1] = {src};
  return par2hip_gf_muladd_region(ctx, d, 1, s, 1, cr, nullptr, 0, len);
}

int32_t par2hip_gf_mul_slice(par2hip_ctx *ctx, uint16_t scalar, uint16_t *data, size_t len) {
  if (!ctx) return PAR2HIP_EINVAL;
  if (len == 0 || scalar == 1) return PAR2HIP_OK;
  if (!data) return PAR2HIP_EINVAL;
  if (scalar == 0) {
    memset(data, 0, len * 2);
    return PAR2HIP_OK;
  }
  // data = scalar*data: overwrite form of the contraction with one source
  Guard g(ctx);
  const size_t rb = len * 2;
  void *din, *dout;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchIn, rb, &din))) return rc;
  if ((rc = scratch(ctx, kScratchOut, rb, &dout))) return rc;
  if ((rc = params_begin(ctx, params_size({2, 8, 8})))) return rc;
  void *d_coeff, *d_srcs, *d_dsts;
  uint16_t *h_coeff = (uint16_t *)params_alloc(ctx, 2, &d_coeff);
  const uint16_t **h_srcs = (const uint16_t **)params_alloc(ctx, 8, &d_srcs);
  uint16_t **h_dsts = (uint16_t **)params_alloc(ctx, 8, &d_dsts);
  h_coeff[0] = scalar;
  h_srcs[0] = (const uint16_t *)din;
  h_dsts[0] = (uint16_t *)dout;
  if ((rc = params_commit(ctx))) return rc;
  P2H_CHECK(ctx, hipMemcpyAsync(din, data, rb, hipMemcpyHostToDevice, ctx->stream));
  GfProblem p;
  p.srcs = (const uint16_t *const *)d_srcs;
  p.dsts = (uint16_t *const *)d_dsts;
  p.coeff = (const uint16_t *)d_coeff;
  p.nsrc = 1;
  p.ndst = 1;
  p.words = len;
  p.accumulate = false;
  p.align_words = 8;
  if ((rc = run_gf(ctx, ctx->stream, p))) return rc;
  P2H_CHECK(ctx, hipMemcpyAsync(data, dout, rb, hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PAR2HIP_OK;
}

// ------------------------------------------------------------- encode --
int32_t par2hip_encode_dev(par2hip_ctx *ctx, const void *d_slices, size_t n, size_t slice_size,
                           const uint32_t *exps, size_t r, void *d_recovery, void *d_md5, void *d_crc) {
  if (!ctx || !d_slices || (r && !d_recovery)) return PAR2HIP_EINVAL;
  Guard g(ctx);
  return encode_body(ctx, (const uint8_t *)d_slices, n, slice_size, exps, r, (uint8_t *)d_recovery,
                     (uint32_t *)d_md5, (uint32_t *)d_crc, 0);
}

int32_t par2hip_encode(par2hip_ctx *ctx, const uint8_t *slices, size_t n, size_t slice_size,
                       const uint32_t *exps, size_t r, uint8_t *recovery, uint8_t *md5_out,
                       uint32_t *crc_out) {
  if (!ctx || !slices || (r && !recovery)) return PAR2HIP_EINVAL;
  Guard g(ctx);
  if (n == 0 || slice_size == 0 || (slice_size & 1))
    return fail(ctx, PAR2HIP_EINVAL, "need n >= 1 and an even, non-zero slice size");
  void *din, *dout, *dh = nullptr;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchIn, n * slice_size, &din))) return rc;
  if ((rc = scratch(ctx, kScratchOut, std::max<size_t>(r, 1) * slice_size, &dout))) return rc;
  if ((md5_out || crc_out) && (rc = scratch(ctx, kScratchHash, n * 20, &dh))) return rc;
  if ((rc = h2d_staged(ctx, din, slices, n * slice_size))) return rc;
  uint32_t *dmd5 = md5_out ? (uint32_t *)dh : nullptr;
  uint32_t *dcrc = crc_out ? (uint32_t *)dh + 4 * n : nullptr;
  if ((rc = encode_body(ctx, (const uint8_t *)din, n, slice_size, exps, r, (uint8_t *)dout, dmd5, dcrc, 0)))
    return rc;
  if (r && (rc = d2h_staged(ctx, recovery, dout, r * slice_size))) return rc;
  if (md5_out) P2H_CHECK(ctx, hipMemcpyAsync(md5_out, dmd5, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  if (crc_out) P2H_CHECK(ctx, hipMemcpyAsync(crc_out, dcrc, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PAR2HIP_OK;
}

// ------------------------------------------------------------- repair --
int32_t par2hip_gf_invert(par2hip_ctx *ctx, const uint16_t *a, size_t m, uint16_t *inv) {
  if (!ctx || (m && (!a || !inv))) return PAR2HIP_EINVAL;
  if (m == 0) return PAR2HIP_OK;
  Guard g(ctx);
  void *aug;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchGj, m * 2 * m * 2, &aug))) return rc;
  std::vector<uint16_t> h(m * 2 * m, 0);
  for (size_t i = 0; i < m; i++) {
    memcpy(&h[i * 2 * m], a + i * m, m * 2);
    h[i * 2 * m + m + i] = 1;
  }
  P2H_CHECK(ctx, hipMemcpyAsync(aug, h.data(), h.size() * 2, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = launch_gauss_jordan(ctx, ctx->stream, (uint16_t *)aug, (int)m, ctx->d_status))) return rc;
  int status = 0;
  P2H_CHECK(ctx, hipMemcpyAsync(&status, ctx->d_status, sizeof status, hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipMemcpyAsync(h.data(), aug, h.size() * 2, hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (status) return fail(ctx, PAR2HIP_ESINGULAR, "Singular matrix");
  for (size_t i = 0; i < m; i++) memcpy(inv + i * m, &h[i * 2 * m + m], m * 2);
  return PAR2HIP_OK;
}

int32_t par2hip_reconstruct_dev(par2hip_ctx *ctx, void *d_slices, size_t n, size_t slice_size,
                                const uint32_t *missing, size_t m, const void *d_recovery,
                                const uint32_t *rec_exps, void *d_out, int32_t out_in_place) {
  if (!ctx || !d_slices || (m && (!missing || !d_recovery || !rec_exps))) return PAR2HIP_EINVAL;
  if (m && !out_in_place && !d_out) return PAR2HIP_EINVAL;
  Guard g(ctx);
  return reconstruct_dev_impl(ctx, (uint8_t *)d_slices, n, slice_size, missing, m, (const uint8_t *)d_recovery,
                              rec_exps, (uint8_t *)d_out, out_in_place != 0);
}

int32_t par2hip_reconstruct(par2hip_ctx *ctx, const uint8_t *slices, size_t n, size_t slice_size,
                            const uint32_t *missing, size_t m, const uint8_t *recovery,
                            const uint32_t *rec_exps, uint8_t *out) {
  if (!ctx || !slices || (m && (!missing || !recovery || !rec_exps || !out))) return PAR2HIP_EINVAL;
  if (m == 0) return PAR2HIP_OK;
  Guard g(ctx);
  if (n == 0 || slice_size == 0 || (slice_size & 1))
    return fail(ctx, PAR2HIP_EINVAL, "need n >= 1 and an even, non-zero slice size");
  void *din, *drec, *dout;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchIn, n * slice_size, &din))) return rc;
  if ((rc = scratch(ctx, kScratchAux, m * slice_size, &drec))) return rc;
  if ((rc = scratch(ctx, kScratchOut, m * slice_size, &dout))) return rc;
  if ((rc = h2d_staged(ctx, din, slices, n * slice_size))) return rc;
  if ((rc = h2d_staged(ctx, drec, recovery, m * slice_size))) return rc;
  if ((rc = reconstruct_dev_impl(ctx, (uint8_t *)din, n, slice_size, missing, m, (const uint8_t *)drec,
                                 rec_exps, (uint8_t *)dout, false)))
    return rc;
  return d2h_staged(ctx, out, dout, m * slice_size);
}

int32_t par2hip_repair_set_dev(par2hip_ctx *ctx, void *d_slices, size_t n, size_t slice_size,
                               const uint8_t *expected_md5, const uint32_t *expected_crc,
                               const void *d_recovery, const uint32_t *rec_exps, size_t nrec,
                               void *d_out, uint32_t *damaged_out, size_t *ndamaged) {
  if (!ctx || !d_slices || !expected_md5 || !expected_crc || !damaged_out || !ndamaged ||
      (nrec && (!d_recovery || !rec_exps)))
    return PAR2HIP_EINVAL;
  Guard g(ctx);
  *ndamaged = 0;
  return repair_set_impl(ctx, (uint8_t *)d_slices, n, slice_size, expected_md5, expected_crc,
                         (const uint8_t *)d_recovery, rec_exps, nrec, (uint8_t *)d_out, damaged_out, ndamaged);
}

// ------------------------------------------------------------- hashes --
int32_t par2hip_hash_slices_dev(par2hip_ctx *ctx, const void *d_slices, size_t n, size_t slice_size,
                                size_t slice_stride, const uint64_t *lens, void *d_md5, void *d_crc) {
  if (!ctx || (n && !d_slices)) return PAR2HIP_EINVAL;
  if (n == 0) return PAR2HIP_OK;
  Guard g(ctx);
  if (slice_stride == 0) slice_stride = slice_size;
  if (slice_stride < slice_size && !lens) return fail(ctx, PAR2HIP_EINVAL, "stride shorter than the slice size");
  if (d_md5 && slice_size >= kMd5MaxSliceBytes) return fail(ctx, PAR2HIP_EINVAL, "hashed slice of 128 GiB or more");
  const uint64_t *d_lens = nullptr;
  int32_t rc;
  if (lens) {
    for (size_t i = 0; i < n; i++)
      if (lens[i] > slice_stride && i + 1 < n)
        return fail(ctx, PAR2HIP_EINVAL, "slice length beyond its stride");
    if ((rc = params_begin(ctx, params_size({n * 8})))) return rc;
    void *dl;
    uint64_t *hl = (uint64_t *)params_alloc(ctx, n * 8, &dl);
    memcpy(hl, lens, n * 8);
    if ((rc = params_commit(ctx))) return rc;
    d_lens = (const uint64_t *)dl;
  }
  if (d_md5 && (rc = launch_md5(ctx, ctx->stream, (const uint8_t *)d_slices, n, slice_size, slice_stride,
                                d_lens, (uint32_t *)d_md5)))
    return rc;
  if (d_crc && (rc = launch_crc32(ctx, ctx->stream, (const uint8_t *)d_slices, n, slice_size, slice_stride,
                                  d_lens, (uint32_t *)d_crc)))
    return rc;
  return PAR2HIP_OK;
}

int32_t par2hip_hash_slices(par2hip_ctx *ctx, const uint8_t *slices, size_t n, size_t slice_size,
                            size_t slice_stride, const uint64_t *lens, uint8_t *md5_out, uint32_t *crc_out) {
  if (!ctx || (n && !slices)) return PAR2HIP_EINVAL;
  if (n == 0) return PAR2HIP_OK;
  Guard g(ctx);
  if (slice_stride == 0) slice_stride = slice_size;
  // bytes actually present in the caller's buffer
  size_t span = (n - 1) * slice_stride + (lens ? (size_t)lens[n - 1] : slice_size);
  if (lens)
    for (size_t i = 0; i < n; i++) span = std::max(span, i * slice_stride + std::min<size_t>(lens[i], slice_size));
  const size_t dstride = (slice_stride + 15) & ~(size_t)15;
  void *din, *dh;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchIn, n * dstride + 16, &din))) return rc;
  if ((rc = scratch(ctx, kScratchHash, n * 20, &dh))) return rc;
  if (dstride == slice_stride) {
    if ((rc = h2d_staged(ctx, din, slices, span))) return rc;
  } else {
    if (n > 1)
      P2H_CHECK(ctx, hipMemcpy2DAsync(din, dstride, slices, slice_stride, std::min(slice_stride, slice_size), n - 1,
                                      hipMemcpyHostToDevice, ctx->stream));
    const size_t last = lens ? std::min<size_t>(lens[n - 1], slice_size) : slice_size;
    P2H_CHECK(ctx, hipMemcpyAsync((uint8_t *)din + (n - 1) * dstride, slices + (n - 1) * slice_stride,
                                  std::min(last, slice_stride), hipMemcpyHostToDevice, ctx->stream));
  }
  std::vector<uint64_t> lv;
  const uint64_t *lp = lens;
  if (!lens && slice_stride < slice_size) return fail(ctx, PAR2HIP_EINVAL, "stride shorter than the slice size");
  if (lens && dstride != slice_stride) {
    lv.assign(lens, lens + n);
    for (auto &x : lv) x = std::min<uint64_t>(x, std::min(slice_stride, slice_size));
    lp = lv.data();
  }
  uint32_t *dmd5 = (uint32_t *)dh, *dcrc = (uint32_t *)dh + 4 * n;
  if ((rc = par2hip_hash_slices_dev(ctx, din, n, slice_size, dstride, lp, md5_out ? dmd5 : nullptr,
                                    crc_out ? dcrc : nullptr)))
    return rc;
  if (md5_out) P2H_CHECK(ctx, hipMemcpyAsync(md5_out, dmd5, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  if (crc_out) P2H_CHECK(ctx, hipMemcpyAsync(crc_out, dcrc, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PAR2HIP_OK;
}

// ------------------------------------------------------------- tuning --
int32_t par2hip_set_gf_kernel(par2hip_ctx *ctx, int32_t which) {
  if (!ctx || which < 0 || which > 3) return PAR2HIP_EINVAL;
  Guard g(ctx);
  int32_t prev = ctx->gf_kernel;
  ctx->gf_kernel = which;
  return prev;
}

int32_t par2hip_set_autotune(par2hip_ctx *ctx, int32_t mode) {
  if (!ctx || mode < 0 || mode > 1) return PAR2HIP_EINVAL;
  Guard g(ctx);
  const int32_t prev = ctx->autotune ? 1 : 0;
  ctx->autotune = mode == 1;
  return prev;
}

int32_t par2hip_force_tile_shape(par2hip_ctx *ctx, int32_t v, int32_t g, int32_t d) {
  if (!ctx) return PAR2HIP_EINVAL;
  Guard gd(ctx);
  if (v == 0 && g == 0 && d == 0) {
    ctx->forced_shape = -1;
    return PAR2HIP_OK;
  }
  const int i = par2hip::tile_shape_index(v, g, d);
  if (i < 0) return par2hip::fail(ctx, PAR2HIP_EINVAL, "no such tile shape");
  ctx->forced_shape = i;
  return PAR2HIP_OK;
}

int32_t par2hip_autotune_save(const char *path) { return path ? par2hip::tune_table_save(path) : PAR2HIP_EINVAL; }

int32_t par2hip_autotune_load(const char *path, size_t *loaded) {
  if (loaded) *loaded = 0;
  return path ? par2hip::tune_table_load(path, loaded) : PAR2HIP_EINVAL;
}

int32_t par2hip_profile_enable(par2hip_ctx *ctx, int32_t on) {
  if (!ctx) return PAR2HIP_EINVAL;
  Guard g(ctx);
  ctx->profile = on != 0;
  return PAR2HIP_OK;
}

int32_t par2hip_profile_read(par2hip_ctx *ctx, par2hip_kernel_stat *out, size_t max, size_t *count) {
  if (!ctx) return PAR2HIP_EINVAL;
  Guard g(ctx);
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->side));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->rest));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->wide));
  for (hipStream_t q : ctx->narrow)
    if (q) P2H_CHECK(ctx, hipStreamSynchronize(q));
  for (auto &p : ctx->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      ctx->stats[p.stat].launches++;
      ctx->stats[p.stat].total_ms += ms;
    }
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
  }
  ctx->pending.clear();
  if (count) *count = ctx->stats.size();
  for (size_t i = 0; i < ctx->stats.size() && i < max; i++) {
    memset(out[i].name, 0, sizeof out[i].name);
    strncpy(out[i].name, ctx->stats[i].name.c_str(), sizeof out[i].name - 1);
    out[i].launches = ctx->stats[i].launches;
    out[i].total_ms = ctx->stats[i].total_ms;
  }
  ctx->stats.clear();
  return PAR2HIP_OK;
}

}  // extern "C"
