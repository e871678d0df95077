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
// One thread per 16-byte table entry (layout in common.hpp).
__global__ void k_nibble_tables(uint64_t *__restrict__ tabs, const uint16_t *__restrict__ coeff,
                                int nrow, int nsrc, int ldc, const uint16_t *__restrict__ log_tab,
                                const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= gf_tab_entries(nrow, nsrc)) return;
  const long gd = idx >> 6;
  const int d = (int)(gd % nsrc), g = (int)(gd / nsrc);
  reinterpret_cast<u32x4 *>(tabs)[idx] = gf_tab_entry(coeff, ldc, nrow, g, d, (int)(idx & 63), log_tab, exp_tab);
}

// Workgroup w covers words [tile*kTpb*V, +kTpb*V) of output rows
// [16*rowblk, +16).  The 1-D grid is remapped so that the row blocks of one
// word tile run on the same XCD back to back (workgroups are dealt to the 8
// XCDs round-robin), letting them share the tile's input lines in that L2.
template <int V>
__global__ __launch_bounds__(kTpb, 2) void k_gf_lds(const uint16_t *const *__restrict__ srcs, int nsrc,
                                                   uint16_t *const *__restrict__ dsts, int ndst,
                                                   const uint64_t *__restrict__ tabs, long words,
                                                   int accumulate, int ntiles, int nrowblk) {
  __shared__ u32x4 lds[kDch * kGpw * 64];  // [d][g][i*16+v], 32 KiB
  const int L = blockIdx.x;
  const int xcd = L & 7, q = L >> 3;
  const int tiles_per_xcd = (ntiles + 7) >> 3;
  // within an XCD: consecutive q walk the row blocks of one tile, then the next tile
  const int local_tile = q / nrowblk, rowblk = q % nrowblk;
  const int tile = local_tile * 8 + xcd;
  if (local_tile >= tiles_per_xcd || tile >= ntiles) return;
  gf_lds_tile<V>(lds, srcs, nsrc, dsts, ndst, tabs, words, accumulate, tile, rowblk, 2 * V);
}

template <int V>
static int32_t launch_lds_variant(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p,
                                  const uint64_t *tabs) {
  const long per_wg = (long)kTpb * V;
  const int ntiles = (int)((p.words + per_wg - 1) / per_wg);
  const int nrowblk = gf_task_rowblks(p.ndst);
  const int tiles_per_xcd = (ntiles + 7) / 8;
  const long nwg = 8L * tiles_per_xcd * nrowblk;
  if (nwg > 0x7fffffffL) return fail(ctx, PAR2HIP_EINVAL, "GF problem too large for one launch");
  ScopedTiming t(ctx, p.tag ? p.tag : "gf_muladd_lds", s);
  hipLaunchKernelGGL(k_gf_lds<V>, dim3((unsigned)nwg), dim3(kTpb), 0, s, p.srcs, p.nsrc, p.dsts,
                     p.ndst, tabs, (long)p.words, p.accumulate ? 1 : 0, ntiles, nrowblk);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

int32_t run_gf_lds(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p) {
  const size_t tab_bytes = (size_t)gf_tab_entries(p.ndst, p.nsrc) * 16;
  void *tabs = nullptr;
  int32_t rc = scratch(ctx, kScratchTables, tab_bytes, &tabs);
  if (rc) return rc;
  {
    const long total = gf_tab_entries(p.ndst, p.nsrc);
    ScopedTiming t(ctx, "nibble_tables", s);
    hipLaunchKernelGGL(k_nibble_tables, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (uint64_t *)tabs, p.coeff, p.ndst, p.nsrc, p.ldc ? p.ldc : p.nsrc, ctx->d_log, ctx->d_exp);
    P2H_CHECK(ctx, hipGetLastError());
  }
  // words per thread: the widest the alignment allows (per-workgroup table
  // staging is amortised over more words); narrower only when the grid would
  // not even give every resident workgroup slot (2 per CU) one workgroup
  const int nrowblk = gf_task_rowblks(p.ndst);
  const long slots = 2L * ctx->num_cus;
  auto nwg = [&](int v) { return (long)((p.words + (size_t)kTpb * v - 1) / ((size_t)kTpb * v)) * nrowblk; };
  int v = p.align_words % 8 == 0 ? 8 : p.align_words % 4 == 0 ? 4 : p.align_words % 2 == 0 ? 2 : 1;
  while (v > 1 && nwg(v) < slots) v /= 2;
  if (v == 8) return launch_lds_variant<8>(ctx, s, p, (const uint64_t *)tabs);
  if (v == 4) return launch_lds_variant<4>(ctx, s, p, (const uint64_t *)tabs);
  if (v == 2) return launch_lds_variant<2>(ctx, s, p, (const uint64_t *)tabs);
  return launch_lds_variant<1>(ctx, s, p, (const uint64_t *)tabs);
}

int32_t run_gf(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p) {
  if (p.ndst <= 0 || p.words == 0) return PAR2HIP_OK;
  if (p.nsrc <= 0) {
    if (p.accumulate) return PAR2HIP_OK;
    return fail(ctx, PAR2HIP_EINVAL, "contraction with no sources and overwrite semantics");
  }
  // 1 = LDS tables, 2 = bit-sliced MFMA (needs 16-byte aligned operands), 0 = automatic
  const bool mfma_ok = p.align_words >= 8;
  int which = ctx->gf_kernel;
  if (which == 0) which = kAutoGfKernel;
  if (which == 2 && mfma_ok) return run_gf_mfma(ctx, s, p);
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

__global__ void k_tables_batch(const GfTaskDev *__restrict__ tasks, const uint16_t *__restrict__ log_tab,
                               const uint16_t *__restrict__ exp_tab) {
  const GfTaskDev t = tasks[blockIdx.y];
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= gf_tab_entries(t.ndst, t.nsrc)) return;
  const long gd = idx >> 6;
  const int d = (int)(gd % t.nsrc), g = (int)(gd / t.nsrc);
  reinterpret_cast<P2H_GLOBAL u32x4 *>(gptr(t.tabs))[idx] =
      gf_tab_entry(t.coeff, t.ldc, t.ndst, g, d, (int)(idx & 63), log_tab, exp_tab);
}

int32_t launch_tables_batch(par2hip_ctx *ctx, hipStream_t s, const GfTaskDev *d_tasks, int ntasks,
                            long max_entries) {
  if (ntasks == 0 || max_entries == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "nibble_tables", s);
  hipLaunchKernelGGL(k_tables_batch, dim3((unsigned)((max_entries + 255) / 256), (unsigned)ntasks), dim3(256), 0, s,
                     d_tasks, ctx->d_log, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

__global__ void k_reset_cursor(unsigned *cursor) { *cursor = 0u; }

// Persistent: every workgroup takes the next work item from the cursor until
// none is left (every wave reaches the exit because the cursor only grows).
// Items are sorted by cost (sources per task, descending) on the host.
__global__ __launch_bounds__(kTpb, 2) void k_gf_lds_batch(const GfTaskDev *__restrict__ tasks,
                                                         const GfItem *__restrict__ items, int nitems,
                                                         unsigned *__restrict__ cursor) {
  __shared__ u32x4 lds[kDch * kGpw * 64];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(cursor, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();  // all lanes hold `it` before lane 0 may overwrite s_item
    if (it >= nitems) break;
    const GfItem w = items[it];
    const GfTaskDev t = tasks[w.task];
    gf_lds_tile<8>(lds, t.srcs, t.nsrc, t.dsts, t.ndst, t.tabs, t.words, t.accumulate, w.tile, w.rowblk, t.align);
  }
}

int32_t launch_gf_batch(par2hip_ctx *ctx, hipStream_t s, const GfTaskDev *d_tasks, const GfItem *d_items,
                        int nitems, unsigned *d_cursor, const char *tag) {
  if (nitems == 0) return PAR2HIP_OK;
  hipLaunchKernelGGL(k_reset_cursor, dim3(1), dim3(1), 0, s, d_cursor);
  ScopedTiming t(ctx, tag, s);
  const int grid = std::min(nitems, 2 * ctx->num_cus);
  hipLaunchKernelGGL(k_gf_lds_batch, dim3(grid), dim3(kTpb), 0, s, d_tasks, d_items, nitems, d_cursor);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

}  // namespace par2hip
