// gf_kernels.hip -- GF(2^16) coefficient generation and the region
// multiply-accumulate ("contraction") kernels for gfx950.
//
// The contraction computes, for every output row i and word w,
//     dst_i[w] (= or ^=) XOR_j coeff[i][j] * src_j[w]
// which is the encoder's recovery accumulation (src/par2_rs.rs:816-870,
// muladd per (recovery, input) pair) and the repair's region muladd
// (src/galois/region/mod.rs:90-131 driven by src/par2_rs.rs:527-603) done as
// one pass over the inputs with the outputs held in registers.
//
// Variant T ("LDS nibble tables"): every coefficient c is expanded into four
// 16-entry tables T_i[v] = c * (v << 4i); four outputs are packed into one
// 64-bit entry, so one ds_read_b64 yields four partial products.  A 16-entry
// table of 8-byte entries covers all 32 LDS banks exactly once, so any lane
// pattern is conflict-free (equal addresses broadcast).
#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>

#include "gf_device.hpp"

namespace par2hip {

// --------------------------------------------------------------- helpers --
__global__ void k_coeff_from_logs(uint16_t *__restrict__ coeff, const uint32_t *__restrict__ src_log,
                                  const uint32_t *__restrict__ exps, int nrow, int nsrc,
                                  const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nrow * nsrc) return;
  const int r = (int)(idx / nsrc), d = (int)(idx % nsrc);
  // c_d^e = 2^(n_d * e) with e already reduced mod 65535 (as gf_pow does,
  // core.rs:256): n_d * e < 2^32, folded twice into [0, 65535]
  uint32_t l = src_log[d] * exps[r];
  l = (l & 0xFFFFu) + (l >> 16);
  l = (l & 0xFFFFu) + (l >> 16);
  coeff[idx] = exp_tab[l];
}

int32_t launch_coeff_from_logs(par2hip_ctx *ctx, hipStream_t s, uint16_t *coeff,
                               const uint32_t *d_src_log, const uint32_t *d_exps, int nrow,
                               int nsrc) {
  const long total = (long)nrow * nsrc;
  if (total == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "coeff_from_logs", s);
  hipLaunchKernelGGL(k_coeff_from_logs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     coeff, d_src_log, d_exps, nrow, nsrc, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// ------------------------------------------------- variant T: nibble tables --
// One wavefront per (8-row group, source) pair, four pairs per workgroup.
__global__ __launch_bounds__(256) void k_nibble_tables(uint64_t *__restrict__ tabs, const uint16_t *__restrict__ coeff,
                                                       int nrow, int nsrc, int ldc) {
  const long pair = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pair >= (long)gf_tab_groups(nrow) * nsrc) return;
  const int d = (int)(pair % nsrc), g = (int)(pair / nsrc);
  gf_tab_pair(reinterpret_cast<P2H_GLOBAL uint16_t *>(gptr(tabs)) + pair * 512, coeff, ldc, nrow, g, d,
              threadIdx.x & 63);
}

// Workgroup w covers words [tile*kTpb*V, +kTpb*V) of output rows
// [8G*rowblk, +8G).  The 1-D grid is remapped so that the row blocks of one
// word tile run on the same XCD back to back (workgroups are dealt to the 8
// XCDs round-robin), letting them share the tile's input lines in that L2.
template <int V, int G, int D>
__global__ __launch_bounds__(kTpb, 2) void k_gf_lds(const uint16_t *const *__restrict__ srcs, int nsrc,
                                                   uint16_t *const *__restrict__ dsts, int ndst,
                                                   const uint64_t *__restrict__ tabs, long words,
                                                   int accumulate, int ntiles, int nrowblk) {
  __shared__ u32x4 lds[D * G * 64];  // [d][g][i*16+v]
  const int L = blockIdx.x;
  const int xcd = L & 7, q = L >> 3;
  const int tiles_per_xcd = (ntiles + 7) >> 3;
  // within an XCD: consecutive q walk the row blocks of one tile, then the next tile
  const int local_tile = q / nrowblk, rowblk = q % nrowblk;
  const int tile = local_tile * 8 + xcd;
  if (local_tile >= tiles_per_xcd || tile >= ntiles) return;
  gf_lds_tile<V, G, D>(lds, srcs, nsrc, dsts, ndst, tabs, words, accumulate, tile, rowblk, 2 * V);
}

template <int V, int G, int D>
static int32_t launch_lds_variant(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, const uint64_t *tabs) {
  const long per_wg = (long)kTpb * V;
  const int ntiles = (int)((p.words + per_wg - 1) / per_wg);
  const int nrowblk = gf_rowblks(p.ndst, G);
  const int tiles_per_xcd = (ntiles + 7) / 8;
  const long nwg = 8L * tiles_per_xcd * nrowblk;
  if (nwg > 0x7fffffffL) return fail(ctx, PAR2HIP_EINVAL, "GF problem too large for one launch");
  ScopedTiming t(ctx, p.tag ? p.tag : "gf_muladd_lds", s);
  hipLaunchKernelGGL((k_gf_lds<V, G, D>), dim3((unsigned)nwg), dim3(kTpb), 0, s, p.srcs, p.nsrc, p.dsts, p.ndst, tabs,
                     (long)p.words, p.accumulate ? 1 : 0, ntiles, nrowblk);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// ---- tile shapes and their tuning
struct TileShape {
  int v, g, d;
};
// Candidates (words per thread, 8-row groups per workgroup, sources per LDS
// chunk); every V has a G = 1 and a G = 2 shape with D = 16 for the default.
constexpr TileShape kShapes[] = {{8, 2, 16}, {8, 1, 16}, {8, 2, 32}, {4, 4, 16}, {4, 2, 16}, {4, 1, 16}, {4, 2, 32},
                                 {2, 4, 16}, {2, 2, 16}, {2, 1, 16}, {1, 4, 16}, {1, 2, 16}, {1, 1, 16}};
constexpr int kNumShapes = (int)(sizeof kShapes / sizeof kShapes[0]);

static int32_t launch_shape(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, const uint64_t *tabs, int i) {
  switch (i) {
    case 0: return launch_lds_variant<8, 2, 16>(ctx, s, p, tabs);
    case 1: return launch_lds_variant<8, 1, 16>(ctx, s, p, tabs);
    case 2: return launch_lds_variant<8, 2, 32>(ctx, s, p, tabs);
    case 3: return launch_lds_variant<4, 4, 16>(ctx, s, p, tabs);
    case 4: return launch_lds_variant<4, 2, 16>(ctx, s, p, tabs);
    case 5: return launch_lds_variant<4, 1, 16>(ctx, s, p, tabs);
    case 6: return launch_lds_variant<4, 2, 32>(ctx, s, p, tabs);
    case 7: return launch_lds_variant<2, 4, 16>(ctx, s, p, tabs);
    case 8: return launch_lds_variant<2, 2, 16>(ctx, s, p, tabs);
    case 9: return launch_lds_variant<2, 1, 16>(ctx, s, p, tabs);
    case 10: return launch_lds_variant<1, 4, 16>(ctx, s, p, tabs);
    case 11: return launch_lds_variant<1, 2, 16>(ctx, s, p, tabs);
    default: return launch_lds_variant<1, 1, 16>(ctx, s, p, tabs);
  }
}

static int shape_index(int v, int g, int d) {
  for (int i = 0; i < kNumShapes; i++)
    if (kShapes[i].v == v && kShapes[i].g == g && kShapes[i].d == d) return i;
  return -1;
}

// Fixed heuristic: the widest V the alignment allows (table staging is
// amortised over more words), narrower only when the grid would not give
// every resident workgroup slot (2 per CU) one workgroup; one 8-row group per
// workgroup when there are at most 8 rows, two otherwise; 16-source chunks.
static int default_shape(const par2hip_ctx *ctx, const GfProblem &p, int vmax) {
  const int g = p.ndst <= kTabRows ? 1 : 2;
  const long slots = 2L * ctx->num_cus;
  const long nrowblk = gf_rowblks(p.ndst, g);
  int v = vmax;
  while (v > 1 && (long)((p.words + (size_t)kTpb * v - 1) / ((size_t)kTpb * v)) * nrowblk < slots) v /= 2;
  return shape_index(v, g, 16);
}

int tile_shape_index(int v, int g, int d) { return shape_index(v, g, d); }

// Process-wide table of tuned shapes, keyed by the problem's class.
std::mutex g_tune_mu;
std::map<uint64_t, int> g_tuned;

static int bit_length(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }
static uint64_t shape_key(const GfProblem &p, int vmax) {
  const uint64_t rows = (uint64_t)std::min(gf_tab_groups(p.ndst), 255);
  return (uint64_t)vmax | rows << 8 | (uint64_t)bit_length((uint64_t)p.nsrc) << 16 |
         (uint64_t)bit_length(p.words) << 24;
}

void batch_tile_shape(const par2hip_ctx *ctx, int policy, int ndst, int nsrc, long words, int align_bytes, int *v,
                      int *g) {
  const int vmax = align_bytes >= 16 ? 8 : align_bytes >= 8 ? 4 : align_bytes >= 4 ? 2 : 1;
  int idx = -1;
  if (ctx->forced_shape >= 0 && kShapes[ctx->forced_shape].v <= vmax) {
    idx = ctx->forced_shape;
  } else if (policy == kBatchShapeTable) {
    GfProblem p;
    p.ndst = ndst;
    p.nsrc = nsrc;
    p.words = (size_t)words;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(shape_key(p, vmax));
    if (it != g_tuned.end()) idx = it->second;
  } else if (policy == kBatchShapeOneGroup) {
    idx = shape_index(8, 1, 16);
  } else if (policy == kBatchShapeSmallV && words <= 65536 && vmax >= 4) {
    idx = shape_index(4, gf_batch_groups(ndst), 16);
  }
  // the batch kernel has v = 8 and 4 (narrower rows take v = 8 with word
  // accesses, as the default does)
  if (idx < 0 || kShapes[idx].v < 4) {
    *v = 8;
    *g = gf_batch_groups(ndst);
    return;
  }
  *v = kShapes[idx].v;
  *g = std::min(kShapes[idx].g, 2);
}

// Times every candidate the alignment allows on the problem itself (the
// fastest of three launches after a warm-up; the outputs are overwritten with
// the same values each time).  A shape replaces the default only when it is
// at least 3% faster, so near-ties keep the heuristic's choice.
static int32_t tune_shape(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, const uint64_t *tabs, int vmax,
                          int *best) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  P2H_CHECK(ctx, hipEventCreate(&e0));
  P2H_CHECK(ctx, hipEventCreate(&e1));
  const bool prof = ctx->profile;
  ctx->profile = false;
  int32_t rc = PAR2HIP_OK;
  const int def = default_shape(ctx, p, vmax);
  auto time_shape = [&](int i, float *ms) -> int32_t {
    int32_t r;
    if ((r = launch_shape(ctx, s, p, tabs, i))) return r;
    *ms = 1e30f;
    for (int k = 0; k < 3; k++) {
      (void)hipEventRecord(e0, s);
      if ((r = launch_shape(ctx, s, p, tabs, i))) return r;
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) return fail(ctx, PAR2HIP_EHIP, "tuning timing");
      float t = 0;
      (void)hipEventElapsedTime(&t, e0, e1);
      *ms = std::min(*ms, t);
    }
    return PAR2HIP_OK;
  };
  float best_ms = 0, def_ms = 0;
  rc = time_shape(def, &def_ms);
  *best = def;
  best_ms = def_ms;
  for (int i = 0; i < kNumShapes && rc == PAR2HIP_OK; i++) {
    if (i == def || kShapes[i].v > vmax) continue;
    float ms = 0;
    if ((rc = time_shape(i, &ms))) break;
    if (ms < best_ms && ms < 0.97f * def_ms) {
      best_ms = ms;
      *best = i;
    }
  }
  ctx->profile = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

static int32_t build_nibble_tables(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, uint64_t *tabs) {
  const long pairs = (long)gf_tab_groups(p.ndst) * p.nsrc;
  ScopedTiming t(ctx, "nibble_tables", s);
  hipLaunchKernelGGL(k_nibble_tables, dim3((unsigned)((pairs + 3) / 4)), dim3(256), 0, s, tabs, p.coeff, p.ndst,
                     p.nsrc, p.ldc ? p.ldc : p.nsrc);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

int32_t run_gf_lds_fixed(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, uint64_t *tabs) {
  if (p.ndst <= 0 || p.words == 0 || p.nsrc <= 0) return PAR2HIP_OK;
  int32_t rc = build_nibble_tables(ctx, s, p, tabs);
  if (rc) return rc;
  const int vmax = p.align_words % 8 == 0 ? 8 : p.align_words % 4 == 0 ? 4 : p.align_words % 2 == 0 ? 2 : 1;
  const int shape = ctx->forced_shape >= 0 && kShapes[ctx->forced_shape].v <= vmax ? ctx->forced_shape
                                                                                    : default_shape(ctx, p, vmax);
  return launch_shape(ctx, s, p, tabs, shape);
}

int32_t run_gf_lds(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p) {
  const size_t tab_bytes = (size_t)gf_tab_entries(p.ndst, p.nsrc) * 16;
  void *tabs = nullptr;
  int32_t rc = scratch(ctx, kScratchTables, tab_bytes, &tabs);
  if (rc) return rc;
  if ((rc = build_nibble_tables(ctx, s, p, (uint64_t *)tabs))) return rc;
  const int vmax = p.align_words % 8 == 0 ? 8 : p.align_words % 4 == 0 ? 4 : p.align_words % 2 == 0 ? 2 : 1;
  int shape = ctx->forced_shape >= 0 && kShapes[ctx->forced_shape].v <= vmax ? ctx->forced_shape : -1;
  if (shape < 0 && ctx->autotune) {
    const uint64_t key = shape_key(p, vmax);
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) shape = it->second;
    }
    if (shape < 0 && !p.accumulate) {  // an accumulating contraction cannot be re-run
      if ((rc = tune_shape(ctx, s, p, (const uint64_t *)tabs, vmax, &shape))) return rc;
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tuned[key] = shape;
    }
  }
  if (shape < 0) shape = default_shape(ctx, p, vmax);
  return launch_shape(ctx, s, p, (const uint64_t *)tabs, shape);
}

int32_t tune_table_save(const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) return PAR2HIP_EINVAL;
  std::lock_guard<std::mutex> g(g_tune_mu);
  fprintf(f, "# par2hip contraction tile shapes: key v g d\n");
  for (const auto &kv : g_tuned)
    fprintf(f, "%llx %d %d %d\n", (unsigned long long)kv.first, kShapes[kv.second].v, kShapes[kv.second].g,
            kShapes[kv.second].d);
  return fclose(f) == 0 ? PAR2HIP_OK : PAR2HIP_EINVAL;
}

int32_t tune_table_load(const char *path, size_t *loaded) {
  FILE *f = fopen(path, "r");
  if (!f) return PAR2HIP_EINVAL;
  char line[256];
  size_t n = 0;
  int32_t rc = PAR2HIP_OK;
  std::lock_guard<std::mutex> g(g_tune_mu);
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    unsigned long long key;
    int v, gg, d;
    if (sscanf(line, "%llx %d %d %d", &key, &v, &gg, &d) != 4 || shape_index(v, gg, d) < 0 || v > (int)(key & 0xff)) {
      rc = PAR2HIP_EINVAL;  // malformed line or an unknown shape: stop, keep what was read
      break;
    }
    g_tuned[(uint64_t)key] = shape_index(v, gg, d);
    n++;
  }
  fclose(f);
  if (loaded) *loaded = n;
  return rc;
}

int32_t run_gf(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p) {
  if (p.ndst <= 0 || p.words == 0) return PAR2HIP_OK;
  if (p.nsrc <= 0) {
    if (p.accumulate) return PAR2HIP_OK;
    return fail(ctx, PAR2HIP_EINVAL, "contraction with no sources and overwrite semantics");
  }
  // The per-coefficient operands (nibble tables: 128 B per coefficient; packed
  // bit matrices: 32 B) grow as ndst * nsrc; beyond kGfChunkBytes of them the
  // output rows are taken in chunks (each an independent contraction over the
  // same sources), so a large inverse or fold never needs more.
  const long per_row = 128L * p.nsrc;
  const long chunk_rows = std::max(64L, (kGfChunkBytes / per_row) & ~63L);
  if (p.ndst > chunk_rows) {
    const int ldc = p.ldc ? p.ldc : p.nsrc;
    for (long r0 = 0; r0 < p.ndst; r0 += chunk_rows) {
      GfProblem c = p;
      c.dsts = p.dsts + r0;
      c.coeff = p.coeff + r0 * ldc;
      c.ldc = ldc;
      c.ndst = (int)std::min<long>(chunk_rows, p.ndst - r0);
      int32_t rc = run_gf(ctx, s, c);
      if (rc) return rc;
    }
    return PAR2HIP_OK;
  }
  // 1 = LDS tables, 2 / 3 = bit-sliced MFMA with FP4 / int8 operands (needs
  // 16-byte aligned operands), 0 = automatic
  const bool mfma_ok = p.align_words >= 8;
  int which = p.kernel ? p.kernel : ctx->gf_kernel;
  // automatic: the LDS tables, except for at most 4 outputs, where the FP4
  // form's 4-output tile reads every input word once and multiplies no
  // padded rows (1 GiB of sources into 4 outputs: 0.26 against 0.42 ms)
  if (which == 0) which = (p.ndst <= 4 && mfma_ok) ? 2 : kAutoGfKernel;
  if ((which == 2 || which == 3) && mfma_ok) return run_gf_mfma(ctx, s, p, which == 3);
  return run_gf_lds(ctx, s, p);
}

// ---------------------------------------------------------------- batched --
__global__ void k_coeff_batch(const CoeffJob *__restrict__ jobs, const uint16_t *__restrict__ exp_tab) {
  const CoeffJob j = jobs[blockIdx.y];
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)j.nrow * j.nsrc) return;
  const int r = (int)(idx / j.nsrc), d = (int)(idx % j.nsrc);
  uint32_t l = j.src_log[d] * j.exps[r];  // exponents arrive reduced mod 65535
  l = (l & 0xFFFFu) + (l >> 16);
  l = (l & 0xFFFFu) + (l >> 16);
  j.coeff[idx] = exp_tab[l];
}

int32_t launch_coeff_batch(par2hip_ctx *ctx, hipStream_t s, const CoeffJob *d_jobs, int njobs, long max_entries) {
  if (njobs == 0 || max_entries == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "coeff_from_logs", s);
  hipLaunchKernelGGL(k_coeff_batch, dim3((unsigned)((max_entries + 255) / 256), (unsigned)njobs), dim3(256), 0, s,
                     d_jobs, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// Every task's tables: pairs (8-row group, source) of all tasks numbered
// consecutively (task t owns [pair0, pair0 + groups*nsrc)), one wavefront
// per pair, grid-stride over the numbering; the owning task is found by
// binary search on pair0 (uniform per wavefront).
__global__ __launch_bounds__(256) void k_tables_batch(const GfTaskDev *__restrict__ tasks, int ntasks, long npairs) {
  const int lane = threadIdx.x & 63;
  const long step = (long)gridDim.x * 4;
  for (long pair = (long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); pair < npairs;
       pair += step) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].pair0 <= pair) lo = mid;
      else hi = mid - 1;
    }
    const GfTaskDev &t = tasks[lo];
    const long q = pair - t.pair0;
    const int d = (int)(q % t.nsrc), g = (int)(q / t.nsrc);
    gf_tab_pair(reinterpret_cast<P2H_GLOBAL uint16_t *>(gptr(t.tabs)) + q * 512, t.coeff, t.ldc, t.ndst, g, d, lane);
  }
}

int32_t launch_tables_batch(par2hip_ctx *ctx, hipStream_t s, const GfTaskDev *d_tasks, int ntasks, long npairs) {
  if (ntasks == 0 || npairs == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "nibble_tables", s);
  const long grid = std::min<long>((npairs + 3) / 4, 16L * ctx->num_cus);
  hipLaunchKernelGGL(k_tables_batch, dim3((unsigned)grid), dim3(256), 0, s, d_tasks, ntasks, npairs);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

__global__ void k_reset_cursor(unsigned *cursor) { *cursor = 0u; }
static_assert(kBatchTpb == kTpb, "batch work items are planned for kTpb threads");

// Persistent: every workgroup takes the next work item from the cursor until
// none is left (every wave reaches the exit because the cursor only grows).
// Items are sorted by cost (sources per task, descending) on the host.
__global__ __launch_bounds__(kTpb, 4) void k_gf_lds_batch(const GfTaskDev *__restrict__ tasks,
                                                         const GfItem *__restrict__ items, int nitems,
                                                         unsigned *__restrict__ cursor) {
  __shared__ u32x4 lds[16 * 2 * 64];  // D = 16 sources x up to 2 groups
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(cursor, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();  // all lanes hold `it` before lane 0 may overwrite s_item
    if (it >= nitems) break;
    const GfItem w = items[it];
    const GfTaskDev t = tasks[w.task];
#define P2H_BATCH_TILE(V, G)                                                                                  \
  gf_lds_tile<V, G, 16, true>(lds, t.srcs, t.nsrc, t.dsts, t.ndst, t.tabs, t.words, t.accumulate, w.tile, w.rowblk, \
                              t.align)
    // the task's tile shape (uniform per item, so the branch is too)
    switch (t.v * 4 + t.groups) {
      case 8 * 4 + 1: P2H_BATCH_TILE(8, 1); break;
      case 4 * 4 + 2: P2H_BATCH_TILE(4, 2); break;
      case 4 * 4 + 1: P2H_BATCH_TILE(4, 1); break;
      default: P2H_BATCH_TILE(8, 2); break;
    }
#undef P2H_BATCH_TILE
  }
}

int32_t launch_gf_batch(par2hip_ctx *ctx, hipStream_t s, const GfTaskDev *d_tasks, const GfItem *d_items,
                        int nitems, unsigned *d_cursor, const char *tag) {
  if (nitems == 0) return PAR2HIP_OK;
  hipLaunchKernelGGL(k_reset_cursor, dim3(1), dim3(1), 0, s, d_cursor);
  // as many persistent workgroups as fit on the device at once (registers and
  // LDS decide; the work list absorbs any that start late)
  static const int per_cu = [] {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(k_gf_lds_batch), kTpb, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 2;
    return nb;
  }();
  ScopedTiming t(ctx, tag, s);
  const int grid = std::min(nitems, per_cu * ctx->num_cus);
  hipLaunchKernelGGL(k_gf_lds_batch, dim3(grid), dim3(kTpb), 0, s, d_tasks, d_items, nitems, d_cursor);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

}  // namespace par2hip
