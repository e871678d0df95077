// noalloc.hip -- par2hip_encode_dev_noalloc: the whole-set encode of
// par2hip_encode_dev with no allocation of any kind during the call, for
// callers that cannot allocate (work under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// Its working memory -- the constant logs and exponents, the row-pointer
// tables, the coefficient matrix and the nibble tables of the LDS-table
// contraction -- lives in one file-scope device array of
// kNoAllocScratchBytes (64 MiB) per device, which each device holds until the
// process ends.  Parameters reach the device as kernel arguments only (the
// exponents in chunks), so the call issues no copy from host memory either.
// Every call on a device shares that device's array: calls that overlap on
// one device (two contexts, or two host threads on different contexts of the
// same device) race; calls queued on one context run in order.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "common.hpp"

// The workspace (device memory, not a host object: host code must use the
// address noalloc_init looks up, never this name).
static __device__ __align__(256) uint8_t scratch[par2hip::kNoAllocScratchBytes];

namespace par2hip {
namespace {

std::mutex g_ws_mu;
std::map<int, void *> g_ws;  // device ordinal -> address of `scratch` on that device

// Number of k in [1, x] coprime to 65535 = 3 * 5 * 17 * 257, by
// inclusion-exclusion over the four prime factors.
__device__ __forceinline__ uint32_t coprime_count(uint32_t x) {
  const uint32_t p[4] = {3, 5, 17, 257};
  int32_t c = 0;
#pragma unroll
  for (int mask = 0; mask < 16; mask++) {
    uint32_t prod = 1;
    int bits = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (mask >> i & 1) {
        prod *= p[i];
        bits++;
      }
    c += (bits & 1 ? -1 : 1) * (int32_t)(x / prod);
  }
  return (uint32_t)c;
}

// log of the d-th PAR2 input constant (par2_rs.rs:64-76): the (d+1)-th n >= 1
// with n mod 3, 5, 17, 257 all non-zero; the 32768 such n lie in [1, 65535].
__device__ __forceinline__ uint32_t constant_log(uint32_t d) {
  uint32_t lo = 1, hi = 65535;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (coprime_count(mid) >= d + 1) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ void k_noalloc_setup(uint32_t *__restrict__ logs, const uint16_t **__restrict__ srcs,
                                uint16_t **__restrict__ dsts, long n, long r, const uint8_t *slices, uint8_t *rec,
                                uint64_t slice_size) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    logs[i] = constant_log((uint32_t)i);
    srcs[i] = reinterpret_cast<const uint16_t *>(slices + (uint64_t)i * slice_size);
  }
  if (i < r) dsts[i] = reinterpret_cast<uint16_t *>(rec + (uint64_t)i * slice_size);
}

// Exponents travel as kernel arguments, kExpChunk per launch.
constexpr int kExpChunk = 896;
struct ExpChunk {
  uint32_t first, count;
  uint32_t e[kExpChunk];
};
__global__ void k_noalloc_exps(uint32_t *__restrict__ exps, ExpChunk c) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c.count) exps[c.first + i] = c.e[i];
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int32_t noalloc_init(par2hip_ctx *ctx) {
  std::lock_guard<std::mutex> g(g_ws_mu);
  auto it = g_ws.find(ctx->device);
  if (it == g_ws.end()) {
    void *p = nullptr;
    // the caller (context creation) has made ctx->device current
    P2H_CHECK(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    it = g_ws.emplace(ctx->device, p).first;
  }
  ctx->noalloc_ws = it->second;
  return PAR2HIP_OK;
}

// Bytes of the workspace an encode of n slices into r recovery slices needs.
static size_t noalloc_bytes(size_t n, size_t r) {
  return up256(n * 4) + up256(r * 4) + up256(n * 8) + up256(r * 8) + up256(r * n * 2) +
         up256((size_t)gf_tab_entries((int)r, (int)n) * 16);
}

}  // namespace par2hip

using namespace par2hip;

extern "C" int32_t par2hip_encode_dev_noalloc(par2hip_ctx *ctx, const void *d_slices, size_t n, size_t slice_size,
                                              const uint32_t *exps, size_t r, void *d_recovery, void *d_md5,
                                              void *d_crc) {
  if (!ctx || !d_slices || (r && !d_recovery)) return PAR2HIP_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (n == 0 || slice_size == 0 || (slice_size & 1))
    return fail(ctx, PAR2HIP_EINVAL, "need n >= 1 and an even, non-zero slice size");
  if (n > 32768) return fail(ctx, PAR2HIP_EINVAL, "more than 32768 input slices repeat PAR2 constants");
  if (r > 65535 || (r && !exps)) return fail(ctx, PAR2HIP_EINVAL, "bad recovery exponent list");
  if (d_md5 && slice_size >= kMd5MaxSliceBytes) return fail(ctx, PAR2HIP_EINVAL, "hashed slice of 128 GiB or more");
  if (noalloc_bytes(n, r) > kNoAllocScratchBytes)
    return fail(ctx, PAR2HIP_EINVAL, "encode needs " + std::to_string(noalloc_bytes(n, r)) +
                                         " bytes of workspace; the allocation-free entry point has " +
                                         std::to_string(kNoAllocScratchBytes));
  if (!ctx->noalloc_ws) return fail(ctx, PAR2HIP_EHIP, "no workspace on this device");
  uint8_t *w = static_cast<uint8_t *>(ctx->noalloc_ws);
  uint32_t *logs = (uint32_t *)w;
  w += up256(n * 4);
  uint32_t *dexps = (uint32_t *)w;
  w += up256(r * 4);
  const uint16_t **srcs = (const uint16_t **)w;
  w += up256(n * 8);
  uint16_t **dsts = (uint16_t **)w;
  w += up256(r * 8);
  uint16_t *coeff = (uint16_t *)w;
  w += up256(r * n * 2);
  uint64_t *tabs = (uint64_t *)w;

  hipStream_t st = ctx->stream;
  const long rows = (long)std::max(n, r);
  hipLaunchKernelGGL(k_noalloc_setup, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, logs, srcs, dsts,
                     (long)n, (long)r, static_cast<const uint8_t *>(d_slices), static_cast<uint8_t *>(d_recovery),
                     (uint64_t)slice_size);
  P2H_CHECK(ctx, hipGetLastError());
  for (size_t first = 0; first < r; first += kExpChunk) {
    ExpChunk c;
    c.first = (uint32_t)first;
    c.count = (uint32_t)std::min<size_t>(kExpChunk, r - first);
    for (uint32_t i = 0; i < c.count; i++) c.e[i] = exps[first + i] % kOrder;  // c^e = c^(e mod 65535)
    hipLaunchKernelGGL(k_noalloc_exps, dim3((c.count + 255) / 256), dim3(256), 0, st, dexps, c);
    P2H_CHECK(ctx, hipGetLastError());
  }
  // MD5 chains on the side CUs, CRC-32 and the contraction on the others,
  // as par2hip_encode_dev does (the contraction here is always the LDS-table
  // form: its only workspace is the nibble tables)
  int32_t rc = PAR2HIP_OK;
  if (d_md5 && (rc = fork_overlap(ctx))) return rc;
  hipStream_t work = d_md5 ? ctx->cur : st;
  if (d_md5 && (rc = launch_md5(ctx, ctx->side, static_cast<const uint8_t *>(d_slices), n, slice_size, slice_size,
                                nullptr, static_cast<uint32_t *>(d_md5), !ctx->masked))) {
    (void)join_overlap(ctx);
    return rc;
  }
  if (d_crc) rc = launch_crc32(ctx, work, static_cast<const uint8_t *>(d_slices), n, slice_size, slice_size, nullptr,
                               static_cast<uint32_t *>(d_crc));
  if (!rc && r) rc = launch_coeff_from_logs(ctx, work, coeff, logs, dexps, (int)r, (int)n);
  if (!rc && r) {
    GfProblem p;
    p.srcs = (const uint16_t *const *)srcs;
    p.dsts = (uint16_t *const *)dsts;
    p.coeff = coeff;
    p.nsrc = (int)n;
    p.ndst = (int)r;
    p.words = slice_size / 2;
    p.accumulate = false;
    p.align_words = std::min({align_words_of((uintptr_t)d_slices), align_words_of((uintptr_t)d_recovery),
                              align_words_of(slice_size)});
    rc = run_gf_lds_fixed(ctx, work, p, tabs);
  }
  if (d_md5) {
    const int32_t jr = join_overlap(ctx);
    if (!rc) rc = jr;
  }
  return rc;
}
