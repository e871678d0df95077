// batch.hip -- many PAR2 sets (encode and repair jobs of uneven sizes) in one
// pass: every job's coefficients, inverses, nibble tables and hashes are
// produced by one launch each over all jobs, and all contractions run in one
// persistent launch over a cost-sorted work list.
#include <algorithm>
#include <cstring>
#include <map>

#include "common.hpp"

namespace par2hip {

// Largest power of two (at most 16) dividing every address in ptrs.
static int row_align(const void *const *ptrs, size_t n) {
  uintptr_t acc = 16;
  for (size_t i = 0; i < n; i++) acc |= (uintptr_t)ptrs[i];
  return (int)(acc & (~acc + 1));
}

namespace {

struct Arena {
  size_t used = 0;
  size_t take(size_t bytes) {
    size_t off = (used + 255) & ~(size_t)255;
    used = off + bytes;
    return off;
  }
};

bool consecutive_exps(const uint32_t *e, size_t m) {
  if (m == 0 || m > (size_t)kVandMaxM) return false;
  for (size_t i = 1; i < m; i++)
    if (e[i] != e[0] + i) return false;
  return e[0] + m - 1 < kOrder;
}

}  // namespace

// Implemented in par2hip_api.hip (the single-set repair path).
int32_t reconstruct_single(par2hip_ctx *ctx, uint8_t *d_slices, size_t n, size_t slice_size, const uint32_t *missing,
                           size_t m, const uint8_t *d_recovery, const uint32_t *rec_exps, uint8_t *d_out);
int32_t fork_overlap(par2hip_ctx *ctx);
int32_t join_overlap(par2hip_ctx *ctx);

size_t batch_job_arena_bytes(const par2hip_job &J) {
  auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t n = J.n, c = J.count;
  if (c == 0) return 0;
  if (J.kind == PAR2HIP_JOB_ENCODE) return r(c * n * 2) + r((size_t)gf_tab_entries((int)c, (int)n) * 16);
  if (!J.exps || !consecutive_exps(J.exps, c)) return 0;  // single-set path after the batch
  const size_t m = c, np = n - m;
  const size_t vs = (std::max<size_t>(np, 1) + 7) & ~(size_t)7, cs = (np + m + 7) & ~(size_t)7;
  size_t t = r(m * 2 * m * 2) + r(m * vs * 2) + r(m * cs * 2);
  if (np) t += r((size_t)gf_tab_entries((int)m, (int)m) * 16);
  return t + r((size_t)gf_tab_entries((int)m, (int)n) * 16);
}

// Everything a batch's launches need, with the parameter block already
// uploaded: built once, launched once (batch_run) or many times (a plan, whose
// parameters and working memory it owns, optionally replayed as a graph).
struct BatchPlan {
  uint8_t *own_params = nullptr, *own_arena = nullptr;
  const GfTaskDev *d_ft = nullptr, *d_mt = nullptr;
  const GfItem *d_fi = nullptr, *d_mi = nullptr;
  const CoeffJob *d_cj = nullptr;
  const InvJob *d_ij = nullptr;
  const HashEntry *d_md5l = nullptr, *d_crcl = nullptr, *d_crcs = nullptr;
  const CrcParams *d_cp = nullptr;
  size_t n_ft = 0, n_mt = 0, n_fi = 0, n_mi = 0, n_cj = 0, n_ij = 0, n_md5 = 0, n_crcl = 0, n_crcs = 0;
  long max_coeff = 0, pairs_ft = 0, pairs_mt = 0;
  int max_m = 0;
  bool aligned4 = true;
  unsigned *cursors = nullptr;
  std::vector<size_t> slow;  // repair jobs for the single-set path (gapped exponents)
};

static int32_t plan_build(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs, uint8_t *ext_arena,
                          size_t ext_bytes, bool own, BatchPlan &P, int policy = kBatchShapeDefault) {
  // ---- validate and classify
  std::vector<size_t> enc, rep, rep_slow;
  for (size_t j = 0; j < njobs; j++) {
    const par2hip_job &J = jobs[j];
    if (!J.d_slices || J.n == 0 || J.slice_size == 0 || (J.slice_size & 1))
      return fail(ctx, PAR2HIP_EINVAL, "batch job: need slices, n >= 1 and an even slice size");
    if (J.n > 32768) return fail(ctx, PAR2HIP_EINVAL, "batch job: more than 32768 input slices");
    // the MD5 kernels count 64-byte blocks in 32 bits
    if (J.slice_size >= kMd5MaxSliceBytes) return fail(ctx, PAR2HIP_EINVAL, "batch job: slice of 128 GiB or more");
    if (J.kind == PAR2HIP_JOB_ENCODE) {
      if (J.count && (!J.exps || !J.d_recovery)) return fail(ctx, PAR2HIP_EINVAL, "batch encode job: exponents and recovery buffer");
      if (J.count > 65535) return fail(ctx, PAR2HIP_EINVAL, "batch encode job: too many recovery slices");
      if (J.count) enc.push_back(j);
    } else if (J.kind == PAR2HIP_JOB_REPAIR) {
      if (J.count == 0) continue;
      if (!J.exps || !J.d_recovery || !J.missing || !J.d_out || J.count > J.n)
        return fail(ctx, PAR2HIP_EINVAL, "batch repair job: exponents, recovery, erasures and output");
      std::vector<char> seen(J.n, 0);
      for (size_t k = 0; k < J.count; k++) {
        if (J.missing[k] >= J.n || seen[J.missing[k]]) return fail(ctx, PAR2HIP_EINVAL, "batch repair job: bad erasure list");
        seen[J.missing[k]] = 1;
      }
      (consecutive_exps(J.exps, J.count) ? rep : rep_slow).push_back(j);
    } else {
      return fail(ctx, PAR2HIP_EINVAL, "batch job: unknown kind");
    }
  }

  // ---- device arena layout
  Arena ar;
  struct EncPlan { size_t coeff; };
  struct RepPlan { size_t aug, vand, coeff; size_t vs, cs, np; };
  std::vector<EncPlan> ep(enc.size());
  std::vector<RepPlan> rp(rep.size());
  for (size_t k = 0; k < enc.size(); k++) {
    const par2hip_job &J = jobs[enc[k]];
    ep[k].coeff = ar.take(J.count * J.n * 2);
  }
  for (size_t k = 0; k < rep.size(); k++) {
    const par2hip_job &J = jobs[rep[k]];
    const size_t m = J.count, np = J.n - m;
    rp[k].np = np;
    rp[k].vs = (std::max<size_t>(np, 1) + 7) & ~(size_t)7;
    rp[k].cs = (np + m + 7) & ~(size_t)7;
    rp[k].aug = ar.take(m * 2 * m * 2);
    rp[k].vand = ar.take(m * rp[k].vs * 2);
    rp[k].coeff = ar.take(m * rp[k].cs * 2);
  }
  // tasks: one fold per repair job (if any present slice), one main per job
  // align: byte alignment all of the task's rows share (rows b + k*s of a
  // span share that of b and s), which bounds the words per thread
  struct TaskPlan { size_t tabs; GfTaskDev t; size_t srcs_at, dsts_at; int align; };
  auto span_align = [](const void *base, size_t stride, size_t count) {
    const void *p[2] = {base, (const uint8_t *)base + stride};
    return row_align(p, count > 1 ? 2 : 1);
  };
  std::vector<TaskPlan> fold_tasks, main_tasks;
  auto tab_bytes = [](int ndst, int nsrc) { return (size_t)gf_tab_entries(ndst, nsrc) * 16; };
  for (size_t k = 0; k < rep.size(); k++) {
    const par2hip_job &J = jobs[rep[k]];
    if (rp[k].np) {
      TaskPlan tp{};
      tp.t.nsrc = (int)J.count;
      tp.t.ndst = (int)J.count;
      tp.tabs = ar.take(tab_bytes(tp.t.ndst, tp.t.nsrc));
      tp.align = 16;  // arena rows: 256-byte aligned blocks, strides of 8 entries
      fold_tasks.push_back(tp);
    }
  }
  for (size_t k = 0; k < enc.size(); k++) {
    const par2hip_job &J = jobs[enc[k]];
    TaskPlan tp{};
    tp.t.nsrc = (int)J.n;
    tp.t.ndst = (int)J.count;
    tp.tabs = ar.take(tab_bytes(tp.t.ndst, tp.t.nsrc));
    tp.align = std::min(span_align(J.d_slices, J.slice_size, J.n), span_align(J.d_recovery, J.slice_size, J.count));
    main_tasks.push_back(tp);
  }
  for (size_t k = 0; k < rep.size(); k++) {
    const par2hip_job &J = jobs[rep[k]];
    TaskPlan tp{};
    tp.t.nsrc = (int)J.n;
    tp.t.ndst = (int)J.count;
    tp.tabs = ar.take(tab_bytes(tp.t.ndst, tp.t.nsrc));
    tp.align = std::min({span_align(J.d_slices, J.slice_size, J.n), span_align(J.d_recovery, J.slice_size, J.count),
                         span_align(J.d_out, J.slice_size, J.count)});
    main_tasks.push_back(tp);
  }
  const size_t cursor_off = ar.take(512);
  void *arena_v = ext_arena;
  int32_t rc;
  if (own) {
    P2H_CHECK(ctx, hipMalloc(&arena_v, std::max<size_t>(ar.used, 4096)));
    P.own_arena = (uint8_t *)arena_v;
  } else if (ext_arena) {
    if (ar.used > ext_bytes) return fail(ctx, PAR2HIP_EINVAL, "batch working memory larger than the arena given");
  } else if ((rc = scratch(ctx, kScratchBatch, std::max<size_t>(ar.used, 4096), &arena_v))) {
    return rc;
  }
  uint8_t *arena = (uint8_t *)arena_v;
  P.slow = rep_slow;

  // ---- work items (sorted by cost: sources per task, descending); each
  // task's tile shape sets its items' word tiles (kTpb * v words) and row
  // blocks (8g rows)
  auto make_items = [&](std::vector<TaskPlan> &tasks, std::vector<GfItem> &items, const std::vector<long> &words) {
    std::vector<int> order(tasks.size());
    for (size_t i = 0; i < tasks.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return (long)tasks[a].t.nsrc * words[a] > (long)tasks[b].t.nsrc * words[b]; });
    for (int ti : order) {
      GfTaskDev &t = tasks[ti].t;
      batch_tile_shape(ctx, policy, t.ndst, t.nsrc, words[ti], tasks[ti].align, &t.v, &t.groups);
      const long per_wg = (long)kBatchTpb * t.v;
      const int ntiles = (int)((words[ti] + per_wg - 1) / per_wg);
      const int nrb = gf_rowblks(t.ndst, t.groups);
      for (int tile = 0; tile < ntiles; tile++)
        for (int rb = 0; rb < nrb; rb++) items.push_back(GfItem{ti, tile, rb, 0});
    }
  };
  std::vector<long> fold_words, main_words;
  {
    size_t f = 0;
    for (size_t k = 0; k < rep.size(); k++)
      if (rp[k].np) {
        fold_words.push_back((long)rp[k].np);
        fold_tasks[f++].t.words = (long)rp[k].np;
      }
    for (size_t k = 0; k < enc.size(); k++) main_words.push_back((long)(jobs[enc[k]].slice_size / 2));
    for (size_t k = 0; k < rep.size(); k++) main_words.push_back((long)(jobs[rep[k]].slice_size / 2));
    for (size_t i = 0; i < main_tasks.size(); i++) main_tasks[i].t.words = main_words[i];
  }
  std::vector<GfItem> fold_items, main_items;
  make_items(fold_tasks, fold_items, fold_words);
  make_items(main_tasks, main_items, main_words);
  if (main_items.size() > 0x7fffffff) return fail(ctx, PAR2HIP_EINVAL, "batch too large");

  // ---- hash lists
  std::vector<HashEntry> md5_list, crc_list, crc_small;
  std::map<uint64_t, uint32_t> crc_idx;
  std::vector<CrcParams> crc_params;
  bool aligned4 = true;
  for (size_t j = 0; j < njobs; j++) {
    const par2hip_job &J = jobs[j];
    if (!J.d_md5 && !J.d_crc) continue;
    uint32_t ci = 0;
    if (J.d_crc) {
      auto it = crc_idx.find(J.slice_size);
      if (it == crc_idx.end()) {
        ci = (uint32_t)crc_params.size();
        crc_idx[J.slice_size] = ci;
        crc_params.push_back(crc_params_for(J.slice_size));
      } else {
        ci = it->second;
      }
    }
    for (uint64_t i = 0; i < J.n; i++) {
      HashEntry e{};
      e.data = (const uint8_t *)J.d_slices + i * J.slice_size;
      e.len = e.size = J.slice_size;
      e.md5 = J.d_md5 ? (uint32_t *)J.d_md5 + 4 * i : nullptr;
      e.crc = J.d_crc ? (uint32_t *)J.d_crc + i : nullptr;
      e.crc_params = ci;
      if (((uintptr_t)e.data & 3) != 0) aligned4 = false;
      if (e.md5) md5_list.push_back(e);
      if (e.crc) (J.slice_size < kCrcLaneMaxBytes ? crc_small : crc_list).push_back(e);
    }
  }
  std::stable_sort(md5_list.begin(), md5_list.end(), [](const HashEntry &a, const HashEntry &b) { return a.size > b.size; });

  // ---- one parameter block for the whole batch
  size_t words_needed = 0;
  for (size_t j : enc) words_needed += (jobs[j].n + jobs[j].count) * 8 + jobs[j].n * 4 + jobs[j].count * 4 + 1024;
  for (size_t k = 0; k < rep.size(); k++) {
    const par2hip_job &J = jobs[rep[k]];
    words_needed += (J.n + 4 * J.count) * 8 + (J.n + J.count) * 4 + 2048;
  }
  const size_t pbytes = words_needed + (fold_tasks.size() + main_tasks.size()) * sizeof(GfTaskDev) +
                        (fold_items.size() + main_items.size()) * sizeof(GfItem) +
                        enc.size() * sizeof(CoeffJob) + rep.size() * sizeof(InvJob) +
                        (md5_list.size() + crc_list.size() + crc_small.size()) * sizeof(HashEntry) +
                        crc_params.size() * sizeof(CrcParams) +
                        64 * 256;
  if (own) P2H_CHECK(ctx, hipMalloc((void **)&P.own_params, pbytes));
  if ((rc = params_begin(ctx, pbytes, P.own_params))) return rc;
  const std::vector<uint32_t> logs_all = constant_logs(32768);
  std::vector<CoeffJob> cjobs;
  long max_coeff = 0;
  for (size_t k = 0; k < enc.size(); k++) {
    const par2hip_job &J = jobs[enc[k]];
    void *d_logs, *d_exps, *d_srcs, *d_dsts;
    uint32_t *h_logs = (uint32_t *)params_alloc(ctx, J.n * 4, &d_logs);
    uint32_t *h_exps = (uint32_t *)params_alloc(ctx, J.count * 4, &d_exps);
    const uint16_t **h_srcs = (const uint16_t **)params_alloc(ctx, J.n * 8, &d_srcs);
    uint16_t **h_dsts = (uint16_t **)params_alloc(ctx, J.count * 8, &d_dsts);
    memcpy(h_logs, logs_all.data(), J.n * 4);
    for (size_t i = 0; i < J.count; i++) h_exps[i] = J.exps[i] % kOrder;
    for (size_t d = 0; d < J.n; d++) h_srcs[d] = (const uint16_t *)((const uint8_t *)J.d_slices + d * J.slice_size);
    for (size_t i = 0; i < J.count; i++) h_dsts[i] = (uint16_t *)((uint8_t *)J.d_recovery + i * J.slice_size);
    uint16_t *coeff = (uint16_t *)(arena + ep[k].coeff);
    cjobs.push_back(CoeffJob{coeff, (const uint32_t *)d_logs, (const uint32_t *)d_exps, (int)J.count, (int)J.n});
    max_coeff = std::max(max_coeff, (long)(J.count * J.n));
    GfTaskDev &t = main_tasks[k].t;
    t.srcs = (const uint16_t *const *)d_srcs;
    t.dsts = (uint16_t *const *)d_dsts;
    t.coeff = coeff;
    t.ldc = (int)J.n;
    t.tabs = (uint64_t *)(arena + main_tasks[k].tabs);
    t.accumulate = 0;
    t.align = std::min(row_align((const void *const *)h_srcs, J.n), row_align((const void *const *)h_dsts, J.count));
  }
  std::vector<InvJob> ijobs;
  int max_m = 0;
  {
    size_t f = 0;
    for (size_t k = 0; k < rep.size(); k++) {
      const par2hip_job &J = jobs[rep[k]];
      const size_t m = J.count, np = rp[k].np;
      std::vector<char> miss(J.n, 0);
      for (size_t q = 0; q < m; q++) miss[J.missing[q]] = 1;
      void *d_mlog, *d_exps, *d_plog, *d_srcs, *d_dsts, *d_fsrc, *d_fdst;
      uint32_t *h_mlog = (uint32_t *)params_alloc(ctx, m * 4, &d_mlog);
      uint32_t *h_exps = (uint32_t *)params_alloc(ctx, m * 4, &d_exps);
      uint32_t *h_plog = (uint32_t *)params_alloc(ctx, std::max<size_t>(np, 1) * 4, &d_plog);
      const uint16_t **h_srcs = (const uint16_t **)params_alloc(ctx, (np + m) * 8, &d_srcs);
      uint16_t **h_dsts = (uint16_t **)params_alloc(ctx, m * 8, &d_dsts);
      const uint16_t **h_fsrc = (const uint16_t **)params_alloc(ctx, m * 8, &d_fsrc);
      uint16_t **h_fdst = (uint16_t **)params_alloc(ctx, m * 8, &d_fdst);
      for (size_t q = 0; q < m; q++) h_mlog[q] = logs_all[J.missing[q]];
      for (size_t i = 0; i < m; i++) h_exps[i] = J.exps[i] % kOrder;
      size_t c = 0;
      for (size_t d = 0; d < J.n; d++)
        if (!miss[d]) {
          h_plog[c] = logs_all[d];
          h_srcs[c] = (const uint16_t *)((const uint8_t *)J.d_slices + d * J.slice_size);
          c++;
        }
      for (size_t i = 0; i < m; i++) h_srcs[np + i] = (const uint16_t *)((const uint8_t *)J.d_recovery + i * J.slice_size);
      for (size_t q = 0; q < m; q++) h_dsts[q] = (uint16_t *)((uint8_t *)J.d_out + q * J.slice_size);
      uint16_t *aug = (uint16_t *)(arena + rp[k].aug);
      uint16_t *vand = (uint16_t *)(arena + rp[k].vand);
      uint16_t *coeff = (uint16_t *)(arena + rp[k].coeff);
      for (size_t i = 0; i < m; i++) {
        h_fsrc[i] = vand + i * rp[k].vs;
        h_fdst[i] = coeff + i * rp[k].cs;
      }
      InvJob ij{};
      ij.aug = aug;
      ij.miss_log = (const uint32_t *)d_mlog;
      ij.plog = (const uint32_t *)d_plog;
      ij.exps = (const uint32_t *)d_exps;
      ij.vand = vand;
      ij.coeff = coeff;
      ij.vs = rp[k].vs;
      ij.cs = rp[k].cs;
      ij.e0 = J.exps[0] % kOrder;
      ij.m = (int)m;
      ij.np = (int)np;
      ijobs.push_back(ij);
      max_m = std::max(max_m, (int)m);
      if (np) {
        GfTaskDev &t = fold_tasks[f].t;
        t.srcs = (const uint16_t *const *)d_fsrc;
        t.dsts = (uint16_t *const *)d_fdst;
        t.coeff = aug + m;  // A^-1, row stride 2m
        t.ldc = (int)(2 * m);
        t.tabs = (uint64_t *)(arena + fold_tasks[f].tabs);
        t.accumulate = 0;
        t.align = std::min(row_align((const void *const *)h_fsrc, m), row_align((const void *const *)h_fdst, m));
        f++;
      }
      GfTaskDev &t = main_tasks[enc.size() + k].t;
      t.srcs = (const uint16_t *const *)d_srcs;
      t.dsts = (uint16_t *const *)d_dsts;
      t.coeff = coeff;
      t.ldc = (int)rp[k].cs;
      t.tabs = (uint64_t *)(arena + main_tasks[enc.size() + k].tabs);
      t.accumulate = 0;
      t.align = std::min(row_align((const void *const *)h_srcs, np + m), row_align((const void *const *)h_dsts, m));
    }
  }
  auto put = [&](const void *src, size_t bytes) -> void * {
    void *dev = nullptr;
    void *h = params_alloc(ctx, std::max<size_t>(bytes, 8), &dev);
    if (bytes) memcpy(h, src, bytes);
    return dev;
  };
  // tasks' (8-row group, source) pairs numbered consecutively for the table build
  std::vector<GfTaskDev> ft(fold_tasks.size()), mt(main_tasks.size());
  long pairs_ft = 0, pairs_mt = 0;
  for (size_t i = 0; i < ft.size(); i++) {
    ft[i] = fold_tasks[i].t;
    ft[i].pair0 = pairs_ft;
    pairs_ft += (long)gf_tab_groups(ft[i].ndst) * ft[i].nsrc;
  }
  for (size_t i = 0; i < mt.size(); i++) {
    mt[i] = main_tasks[i].t;
    mt[i].pair0 = pairs_mt;
    pairs_mt += (long)gf_tab_groups(mt[i].ndst) * mt[i].nsrc;
  }
  const GfTaskDev *d_ft = (const GfTaskDev *)put(ft.data(), ft.size() * sizeof(GfTaskDev));
  const GfTaskDev *d_mt = (const GfTaskDev *)put(mt.data(), mt.size() * sizeof(GfTaskDev));
  const GfItem *d_fi = (const GfItem *)put(fold_items.data(), fold_items.size() * sizeof(GfItem));
  const GfItem *d_mi = (const GfItem *)put(main_items.data(), main_items.size() * sizeof(GfItem));
  const CoeffJob *d_cj = (const CoeffJob *)put(cjobs.data(), cjobs.size() * sizeof(CoeffJob));
  const InvJob *d_ij = (const InvJob *)put(ijobs.data(), ijobs.size() * sizeof(InvJob));
  const HashEntry *d_md5l = (const HashEntry *)put(md5_list.data(), md5_list.size() * sizeof(HashEntry));
  const HashEntry *d_crcl = (const HashEntry *)put(crc_list.data(), crc_list.size() * sizeof(HashEntry));
  const HashEntry *d_crcs = (const HashEntry *)put(crc_small.data(), crc_small.size() * sizeof(HashEntry));
  const CrcParams *d_cp = (const CrcParams *)put(crc_params.data(), crc_params.size() * sizeof(CrcParams));
  if ((rc = params_commit(ctx))) return rc;
  P.d_ft = d_ft;
  P.d_mt = d_mt;
  P.d_fi = d_fi;
  P.d_mi = d_mi;
  P.d_cj = d_cj;
  P.d_ij = d_ij;
  P.d_md5l = d_md5l;
  P.d_crcl = d_crcl;
  P.d_crcs = d_crcs;
  P.d_cp = d_cp;
  P.n_ft = ft.size();
  P.n_mt = mt.size();
  P.n_fi = fold_items.size();
  P.n_mi = main_items.size();
  P.n_cj = cjobs.size();
  P.n_ij = ijobs.size();
  P.n_md5 = md5_list.size();
  P.n_crcl = crc_list.size();
  P.n_crcs = crc_small.size();
  P.max_coeff = max_coeff;
  P.pairs_ft = pairs_ft;
  P.pairs_mt = pairs_mt;
  P.max_m = max_m;
  P.aligned4 = aligned4;
  P.cursors = (unsigned *)(arena + cursor_off);
  return PAR2HIP_OK;
}

// The launches of a planned batch, queued on the ctx streams (the ctx stream
// is the one to wait on; everything forked from it is joined back).  The
// hashes do not feed the contraction, so they overlap it: MD5 chains on the
// reserved side CUs when they fit there (one wave per SIMD) and from the
// start; otherwise -- a throughput load of many chains -- MD5 and the CRC-32
// lists go to an unmasked stream over the whole device.  That stream is
// forked just before the main contraction, so the persistent contraction
// workgroups and the hash workgroups become ready together and share the
// device, instead of the many short hash workgroups taking every slot first.
// Waves of the hash lists find room only beside fewer than four contraction
// workgroups per CU, so most MD5 chains run as the contraction drains; the
// longest chains are listed first.
static int32_t plan_launch(par2hip_ctx *ctx, const BatchPlan &P) {
  int32_t rc;
  const size_t md5_waves = (P.n_md5 + 63) / 64;
  const bool side_fits = ctx->masked && md5_waves <= (size_t)(4 * ctx->hash_cus);
  const bool md5_side = P.n_md5 && side_fits;
  const bool md5_wide = P.n_md5 && !side_fits;
  const bool wide = md5_wide || P.n_crcl || P.n_crcs;
  if (md5_side) {
    if ((rc = fork_overlap(ctx))) return rc;
    if ((rc = launch_md5_list(ctx, ctx->side, P.d_md5l, P.n_md5, P.aligned4, false))) return rc;
  }
  hipStream_t s = ctx->cur;
  if ((rc = launch_coeff_batch(ctx, s, P.d_cj, (int)P.n_cj, P.max_coeff))) return rc;
  if ((rc = launch_inverse_batch(ctx, s, P.d_ij, (int)P.n_ij, P.max_m, 0))) return rc;
  if ((rc = launch_tables_batch(ctx, s, P.d_ft, (int)P.n_ft, P.pairs_ft))) return rc;
  if ((rc = launch_gf_batch(ctx, s, P.d_ft, P.d_fi, (int)P.n_fi, P.cursors, "gf_fold"))) return rc;
  if ((rc = launch_tables_batch(ctx, s, P.d_mt, (int)P.n_mt, P.pairs_mt))) return rc;
  if (wide) {
    P2H_CHECK(ctx, hipEventRecord(ctx->fork, s));
    P2H_CHECK(ctx, hipStreamWaitEvent(ctx->wide, ctx->fork, 0));
    if (md5_wide && (rc = launch_md5_list(ctx, ctx->wide, P.d_md5l, P.n_md5, P.aligned4, false))) return rc;
    if ((rc = launch_crc32_list(ctx, ctx->wide, P.d_crcl, P.n_crcl, P.d_cp))) return rc;
    if ((rc = launch_crc32_list_lane(ctx, ctx->wide, P.d_crcs, P.n_crcs))) return rc;
  }
  if ((rc = launch_gf_batch(ctx, s, P.d_mt, P.d_mi, (int)P.n_mi, P.cursors + 64, "gf_batch"))) return rc;
  if (md5_side && (rc = join_overlap(ctx))) return rc;
  if (wide) {
    P2H_CHECK(ctx, hipEventRecord(ctx->join, ctx->wide));
    P2H_CHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->join, 0));
  }
  return PAR2HIP_OK;
}

int32_t batch_run(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs, uint8_t *ext_arena, size_t ext_bytes) {
  BatchPlan P;
  int32_t rc;
  if ((rc = plan_build(ctx, jobs, njobs, ext_arena, ext_bytes, false, P))) return rc;
  if ((rc = plan_launch(ctx, P))) return rc;
  // repair jobs whose exponents are not consecutive take the Gauss-Jordan path
  for (size_t j : P.slow) {
    const par2hip_job &J = jobs[j];
    if ((rc = reconstruct_single(ctx, (uint8_t *)J.d_slices, J.n, J.slice_size, J.missing, J.count,
                                 (const uint8_t *)J.d_recovery, J.exps, (uint8_t *)J.d_out)))
      return rc;
  }
  return PAR2HIP_OK;
}

}  // namespace par2hip

struct par2hip_batch_plan {
  par2hip_ctx *ctx = nullptr;
  par2hip::BatchPlan plan;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  par2hip_batch_plan_tuning tuning{};
};

static void plan_free(par2hip_batch_plan *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->plan.own_params) (void)hipFree(p->plan.own_params);
  if (p->plan.own_arena) (void)hipFree(p->plan.own_arena);
  delete p;
}

// One plan under one tile-shape policy, captured into a graph when asked.
static int32_t plan_create_one(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs, int32_t use_graph, int policy,
                               par2hip_batch_plan **out) {
  auto *p = new par2hip_batch_plan();
  p->ctx = ctx;
  int32_t rc = par2hip::plan_build(ctx, jobs, njobs, nullptr, 0, true, p->plan, policy);
  if (rc == PAR2HIP_OK && !p->plan.slow.empty())
    rc = par2hip::fail(ctx, PAR2HIP_EINVAL, "a planned batch needs consecutive recovery exponents in every repair job");
  if (rc == PAR2HIP_OK && use_graph) {
    // capture the launches (the parameter upload above is not part of the
    // graph: replays read the plan's own block); timing events stay out
    const bool prof = ctx->profile;
    ctx->profile = false;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      rc = par2hip::plan_launch(ctx, p->plan);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
      p->graph = g;
      if (rc == PAR2HIP_OK && e2 != hipSuccess) rc = par2hip::hip_fail(ctx, e2, "hipStreamEndCapture");
      if (rc == PAR2HIP_OK) {
        e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = par2hip::hip_fail(ctx, e, "hipGraphInstantiate");
      }
    } else {
      rc = par2hip::hip_fail(ctx, e, "hipStreamBeginCapture");
    }
    ctx->cur = ctx->stream;
    ctx->profile = prof;
  }
  if (rc != PAR2HIP_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    plan_free(p);
    return rc;
  }
  p->tuning.policy = policy;
  *out = p;
  return PAR2HIP_OK;
}

// Mean time of `reps` launches after a warm-up one (HIP events on the ctx
// stream; the plan's outputs are rewritten with the same values each time).
static int32_t time_plan(par2hip_batch_plan *p, int reps, double *ms) {
  par2hip_ctx *ctx = p->ctx;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  P2H_CHECK(ctx, hipEventCreate(&e0));
  P2H_CHECK(ctx, hipEventCreate(&e1));
  int32_t rc = par2hip_batch_plan_launch(p);
  (void)hipEventRecord(e0, ctx->stream);
  for (int i = 0; i < reps && rc == PAR2HIP_OK; i++) rc = par2hip_batch_plan_launch(p);
  (void)hipEventRecord(e1, ctx->stream);
  float t = 0;
  if (rc == PAR2HIP_OK && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess)
    *ms = t / reps;
  else if (rc == PAR2HIP_OK)
    rc = par2hip::fail(ctx, PAR2HIP_EHIP, "plan timing");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

extern "C" int32_t par2hip_batch_plan_create(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs,
                                             int32_t use_graph, par2hip_batch_plan **out) {
  if (!ctx || !out || njobs == 0 || !jobs) return PAR2HIP_EINVAL;
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (!ctx->autotune || ctx->forced_shape >= 0) return plan_create_one(ctx, jobs, njobs, use_graph, par2hip::kBatchShapeDefault, out);
  // autotuning: every shape policy planned and timed on the batch itself
  // (the plan is for repeated launches, so its tuning is paid once); the
  // fastest is kept, the default unless another is at least 2% faster
  par2hip_batch_plan *best = nullptr;
  par2hip_batch_plan_tuning info{};
  info.policies = par2hip::kBatchShapePolicies;
  double best_ms = 0, ms0 = 0;
  for (int pol = 0; pol < par2hip::kBatchShapePolicies; pol++) {
    par2hip_batch_plan *p = nullptr;
    int32_t rc = plan_create_one(ctx, jobs, njobs, use_graph, pol, &p);
    double ms = 0;
    if (rc == PAR2HIP_OK) rc = time_plan(p, 2, &ms);
    if (rc != PAR2HIP_OK) {
      if (p) plan_free(p);
      if (best) plan_free(best);
      return rc;
    }
    info.ms[pol] = ms;
    (void)par2hip_sync(ctx);  // nothing of the timed launches is left in flight
    if (pol == 0) {
      best = p;
      best_ms = ms0 = ms;
    } else if (ms < 0.98 * ms0 && ms < best_ms) {
      plan_free(best);
      best = p;
      best_ms = ms;
    } else {
      plan_free(p);
    }
  }
  info.policy = best->tuning.policy;
  best->tuning = info;
  *out = best;
  return PAR2HIP_OK;
}

extern "C" int32_t par2hip_batch_plan_tuning_info(par2hip_batch_plan *p, par2hip_batch_plan_tuning *out) {
  if (!p || !out) return PAR2HIP_EINVAL;
  *out = p->tuning;
  return PAR2HIP_OK;
}

extern "C" int32_t par2hip_batch_plan_launch(par2hip_batch_plan *p) {
  if (!p) return PAR2HIP_EINVAL;
  par2hip_ctx *ctx = p->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (p->exec) {
    P2H_CHECK(ctx, hipGraphLaunch(p->exec, ctx->stream));
    return PAR2HIP_OK;
  }
  return par2hip::plan_launch(ctx, p->plan);
}

extern "C" int32_t par2hip_batch_plan_destroy(par2hip_batch_plan *p) {
  if (!p) return PAR2HIP_EINVAL;
  {
    std::lock_guard<std::recursive_mutex> lk(p->ctx->mu);
    (void)par2hip_sync(p->ctx);
  }
  plan_free(p);
  return PAR2HIP_OK;
}

extern "C" int32_t par2hip_batch_run_dev(par2hip_ctx *ctx, const par2hip_job *jobs, size_t njobs) {
  if (!ctx || (njobs && !jobs)) return PAR2HIP_EINVAL;
  if (njobs == 0) return PAR2HIP_OK;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  return par2hip::batch_run(ctx, jobs, njobs, nullptr, 0);
}
