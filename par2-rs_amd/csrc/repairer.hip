// repairer.hip -- repairing the files of a PAR2 set over the GPU entry
// points: what Par2Repairer does with the reference's verifier and repair
// driver, with the slice hashes (verify and pre-commit check) and the rebuild
// on the device.  Loading the set (parser, volumes, canonical order, block
// layout) is host-only code in par2set.cpp.
//
//   verify  src/verify.rs:87-330, :662-760 (size first; IFSC block checks, or
//           the 16 KiB and whole-file MD5 without IFSC; a missing file or a
//           size mismatch makes every block of the file damaged)
//   repair  src/repair.rs:445-1125 (lowest exponents first, <final>.par2tmp
//           seeded with the original, rebuilt blocks written, every temp
//           verified, renamed only when all verify, temps removed otherwise)
//   cancel  src/lib.rs:196-345, src/repair.rs:56-59, :808-813, :972-977,
//           :1043-1067 (a flag checked at safe points; before the renames
//           only temps exist, so a cancelled call leaves every original
//           byte-identical)
#include <dirent.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "md5_host.hpp"
#include "par2set.hpp"

namespace par2hip {
namespace {

const char kTmpSuffix[] = ".par2tmp";  // src/repair.rs:63-67

bool file_size(const std::string &p, uint64_t *sz) {
  struct stat st;
  if (stat(p.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return false;
  *sz = (uint64_t)st.st_size;
  return true;
}

// Byte ranges of files moved with several host threads (pread / pwrite in
// 8 MiB pieces for reads, one thread per file for writes): reading a set's
// files and writing its temps from one thread leaves most of the host's
// memory bandwidth unused.
struct IoRange {
  std::string path;
  uint64_t off, n;
  uint8_t *buf;
};

bool io_parallel(const std::vector<IoRange> &jobs, bool write) {
  struct Piece {
    size_t job;
    uint64_t off, n;
  };
  std::vector<Piece> pieces;
  const uint64_t kPiece = 8u << 20;
  for (size_t j = 0; j < jobs.size(); j++) {
    if (write || jobs[j].n == 0) {
      pieces.push_back({j, 0, jobs[j].n});
      continue;
    }
    for (uint64_t o = 0; o < jobs[j].n; o += kPiece) pieces.push_back({j, o, std::min(kPiece, jobs[j].n - o)});
  }
  std::atomic<size_t> next{0};
  std::atomic<bool> ok{true};
  auto worker = [&] {
    for (size_t i; ok && (i = next++) < pieces.size();) {
      const Piece &pc = pieces[i];
      const IoRange &jb = jobs[pc.job];
      const int fd = write ? open(jb.path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0666) : open(jb.path.c_str(), O_RDONLY);
      if (fd < 0) {
        ok = false;
        break;
      }
      uint64_t done = 0;
      while (done < pc.n) {
        const ssize_t r = write ? pwrite(fd, jb.buf + pc.off + done, pc.n - done, (off_t)(jb.off + pc.off + done))
                                : pread(fd, jb.buf + pc.off + done, pc.n - done, (off_t)(jb.off + pc.off + done));
        if (r <= 0) break;
        done += (uint64_t)r;
      }
      // start the writeback here, in parallel: ext4 otherwise forces the
      // block allocation of a file renamed over an existing one inside the
      // rename (one file after another under the directory lock)
      if (write && done == pc.n) (void)sync_file_range(fd, 0, 0, SYNC_FILE_RANGE_WRITE);
      if (close(fd) != 0 || done != pc.n) ok = false;
    }
  };
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({pieces.size(), (size_t)16, (size_t)hw});
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; t++) th.emplace_back(worker);
  worker();
  for (auto &t : th) t.join();
  return ok;
}


bool md5_file_prefix(const std::string &path, uint64_t n, uint8_t out[16]) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  Md5Host h;
  std::vector<uint8_t> buf(1 << 20);
  uint64_t left = n;
  bool ok = true;
  while (left) {
    const size_t want = (size_t)std::min<uint64_t>(left, buf.size());
    if (fread(buf.data(), 1, want, f) != want) {
      ok = false;
      break;
    }
    h.update(buf.data(), want);
    left -= want;
  }
  fclose(f);
  h.finish(out);
  return ok;
}

}  // namespace

int32_t set_info(const char *index, par2hip_set_info *info) {
  Par2Set s;
  std::string err;
  int32_t rc = load_par2_set(index, s, &err);
  if (rc) return rc;
  memset(info, 0, sizeof *info);
  info->block_size = s.block_size;
  info->files = (uint32_t)s.files.size();
  memcpy(info->set_id, s.set_id.data(), 16);
  for (const FileRec &r : s.files_in_order) {
    const uint64_t k = blocks_of(r.length, s.block_size);
    info->blocks = k > UINT64_MAX - info->blocks ? UINT64_MAX : info->blocks + k;  // saturates
  }
  const std::vector<RecSlice> recs = recovery_by_exponent(s);
  info->recovery_slices = (uint32_t)recs.size();
  if (!recs.empty()) {
    info->min_exponent = recs.front().exponent;
    info->max_exponent = recs.back().exponent;
  }
  return PAR2HIP_OK;
}

// Relaxed read of the caller's cancel flag (a Rust AtomicBool has the layout
// of a bool: one byte, 0 or 1).
bool cancelled(const uint8_t *flag) { return flag && __atomic_load_n(flag, __ATOMIC_RELAXED) != 0; }

int32_t repair_files(par2hip_ctx *ctx, const char *index, const char *base_dir, int verify_only,
                     const uint8_t *cancel, par2hip_repair_report *rep) {
  memset(rep, 0, sizeof *rep);
  if (cancelled(cancel)) return fail(ctx, PAR2HIP_ECANCELLED, "cancelled");
  Par2Set s;
  std::string err;
  int32_t rc = load_par2_set(index, s, &err);
  if (rc) return fail(ctx, rc, err);
  const std::vector<FileRec> files = canonical_files(s);
  for (const FileRec &f : files)
    if (!safe_name(f.name)) return fail(ctx, PAR2HIP_EINVAL, "unsafe file name in the set: " + f.name);
  const uint64_t bs = s.block_size;
  const std::string base = base_dir && *base_dir ? base_dir : dir_of(index);

  // ---- block layout of the files in canonical order (repair.rs:499-545),
  // refused before any buffer is sized from it when the block count passes
  // the 32768 distinct PAR2 constants or the blocks do not fit the device
  const std::vector<RecSlice> recs = recovery_by_exponent(s);
  rep->files = (uint32_t)files.size();
  rep->block_size = bs;
  rep->recovery_slices = (uint32_t)recs.size();
  size_t dev_free = 0, dev_total = 0;
  P2H_CHECK(ctx, hipMemGetInfo(&dev_free, &dev_total));
  BlockLayout lay;
  if ((rc = block_layout(files, bs, 32768, dev_total, &lay, &err))) return fail(ctx, rc, err);
  const std::vector<uint64_t> &first = lay.first, &nblk = lay.nblk;
  const uint64_t n = lay.n;
  rep->blocks = n;
  if (n + recs.size() > 65535) return fail(ctx, PAR2HIP_EINVAL, "data + recovery blocks exceed 65535");
  if (n == 0) return PAR2HIP_OK;
  if (cancelled(cancel)) return fail(ctx, PAR2HIP_ECANCELLED, "cancelled");

  // ---- verify (verify.rs:87-330): every file's blocks on the device
  enum { kOk, kDamaged, kMissing, kSize };
  std::vector<int> status(files.size(), kOk);
  std::vector<uint64_t> present_size(files.size(), 0);
  // the files' blocks, zero-padded, in one pinned staging copy of the
  // device layout, read by parallel threads and shipped in one copy
  void *d_slices = nullptr, *d_hash = nullptr;
  uint8_t *stage = nullptr;
  if ((rc = scratch(ctx, kScratchIn, n * bs, &d_slices))) return rc;
  if ((rc = scratch(ctx, kScratchHash, n * 20, &d_hash))) return rc;
  if ((rc = host_stage(ctx, n * bs, &stage))) return rc;
  std::vector<IoRange> reads;
  for (size_t i = 0; i < files.size(); i++) {
    const std::string path = base + "/" + files[i].name;
    uint8_t *dst = stage + first[i] * bs;
    uint64_t have = 0;
    if (!file_size(path, &present_size[i])) {
      status[i] = kMissing;
    } else if (present_size[i] != files[i].length) {  // verify_file_smart: size first
      status[i] = kSize;
    } else {
      have = files[i].length;
      if (have) reads.push_back({path, 0, have, dst});
    }
    memset(dst + have, 0, nblk[i] * bs - have);
  }
  if (!io_parallel(reads, false)) return fail(ctx, PAR2HIP_EINVAL, "cannot read the protected files");
  P2H_CHECK(ctx, hipMemcpyAsync(d_slices, stage, n * bs, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = par2hip_hash_slices_dev(ctx, d_slices, n, bs, bs, nullptr, d_hash, (uint8_t *)d_hash + n * 16)))
    return rc;
  std::vector<uint8_t> md5s(n * 16);
  std::vector<uint32_t> crcs(n);
  P2H_CHECK(ctx, hipMemcpyAsync(md5s.data(), d_hash, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  P2H_CHECK(ctx, hipMemcpyAsync(crcs.data(), (uint8_t *)d_hash + n * 16, n * 4, hipMemcpyDeviceToHost,
                                ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<uint32_t> missing;  // global block indices, canonical order
  for (size_t i = 0; i < files.size(); i++) {
    bool all = status[i] != kOk;
    std::vector<uint32_t> bad;
    if (!all) {
      auto it = s.ifsc.find(files[i].id);
      if (it != s.ifsc.end()) {  // IFSC block checks (verify.rs:548-654)
        for (uint64_t k = 0; k < nblk[i]; k++) {
          const uint64_t g = first[i] + k;
          const bool known = k < it->second.size();
          if (!known || crcs[g] != it->second[k].second || memcmp(&md5s[g * 16], it->second[k].first.data(), 16) != 0)
            bad.push_back((uint32_t)g);
        }
      } else {  // 16 KiB then whole-file MD5 (verify.rs:700-745)
        uint8_t d[16];
        const std::string path = base + "/" + files[i].name;
        const bool ok16 = md5_file_prefix(path, std::min<uint64_t>(files[i].length, 16384), d) &&
                          (files[i].length < 16384 || memcmp(d, files[i].md5k.data(), 16) == 0);
        all = !(ok16 && md5_file_prefix(path, files[i].length, d) && memcmp(d, files[i].md5.data(), 16) == 0);
      }
    }
    if (all)
      for (uint64_t k = 0; k < nblk[i]; k++) bad.push_back((uint32_t)(first[i] + k));
    if (!bad.empty() && status[i] == kOk) status[i] = kDamaged;
    missing.insert(missing.end(), bad.begin(), bad.end());
    if (status[i] == kMissing) rep->files_missing++;
    else if (status[i] != kOk) rep->files_damaged++;
  }
  rep->blocks_damaged = missing.size();
  if (cancelled(cancel)) return fail(ctx, PAR2HIP_ECANCELLED, "cancelled");  // lib.rs:246-249, :339-343
  if (missing.empty() || verify_only) return PAR2HIP_OK;
  const size_t m = missing.size();
  if (m > recs.size())
    return fail(ctx, PAR2HIP_ERECOVERY,
                "need " + std::to_string(m) + " recovery blocks, have " + std::to_string(recs.size()));

  // ---- rebuild on the device from the m lowest exponents (repair.rs:575-604);
  // the data is on the device, so the staging copy now takes the recovery
  // slices (m <= n)
  std::vector<uint32_t> exps(m);
  std::vector<IoRange> rreads;
  for (size_t k = 0; k < m; k++) {
    exps[k] = recs[k].exponent;
    rreads.push_back({recs[k].path, recs[k].data_offset, bs, stage + k * bs});
  }
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // the staging copy of the data has left
  if (!io_parallel(rreads, false)) return fail(ctx, PAR2HIP_EINVAL, "cannot read the recovery slices");
  void *d_rec = nullptr, *d_out = nullptr;
  if ((rc = scratch(ctx, kScratchAux, m * bs, &d_rec))) return rc;
  if ((rc = scratch(ctx, kScratchOut, m * bs, &d_out))) return rc;
  P2H_CHECK(ctx, hipMemcpyAsync(d_rec, stage, m * bs, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = par2hip_reconstruct_dev(ctx, d_slices, n, bs, missing.data(), m, d_rec, exps.data(), d_out, 0)))
    return rc;
  if (cancelled(cancel)) {  // repair.rs:972-977: nothing written yet
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return fail(ctx, PAR2HIP_ECANCELLED, "cancelled");
  }

  // ---- temp files (repair.rs:736-1040): originals untouched until commit.
  // Each damaged file's full contents are assembled on the device (its
  // blocks as read, with the rebuilt ones copied over them) and written in
  // one piece; the reference seeds the temp with the original's bytes and
  // writes the rebuilt blocks into it, which gives the same file.
  std::vector<std::pair<std::string, std::string>> temps;  // tmp, final
  auto cleanup = [&] {
    for (auto &t : temps) (void)remove(t.first.c_str());
  };
  std::vector<size_t> touched;
  std::vector<IoRange> writes;
  for (size_t i = 0; i < files.size(); i++) {
    if (status[i] == kOk) continue;
    touched.push_back(i);
    const std::string final_path = base + "/" + files[i].name;
    temps.push_back({final_path + kTmpSuffix, final_path});
    writes.push_back({temps.back().first, 0, files[i].length, stage + first[i] * bs});
  }
  for (size_t k = 0; k < m; k++)
    P2H_CHECK(ctx, hipMemcpyAsync((uint8_t *)d_slices + (uint64_t)missing[k] * bs, (const uint8_t *)d_out + k * bs,
                                  bs, hipMemcpyDeviceToDevice, ctx->stream));
  for (size_t i : touched)
    if (files[i].length)
      P2H_CHECK(ctx, hipMemcpyAsync(stage + first[i] * bs, (const uint8_t *)d_slices + first[i] * bs, files[i].length,
                                    hipMemcpyDeviceToHost, ctx->stream));
  P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (!io_parallel(writes, true)) {
    cleanup();
    return fail(ctx, PAR2HIP_EINVAL, "cannot write the temp files");
  }
  if (cancelled(cancel)) {  // repair.rs:1043-1047: only temps exist
    cleanup();
    return fail(ctx, PAR2HIP_ECANCELLED, "cancelled");
  }

  // ---- every temp verified before anything is renamed (repair.rs:1050-1080):
  // read back from disk (parallel) into the set's device layout, then one
  // hash pass over every block (the chains run in parallel, so the pass takes
  // one chain's time however many files were rewritten)
  std::vector<IoRange> vreads;
  bool any_ifsc = false;
  for (size_t t = 0; t < touched.size(); t++) {
    const size_t i = touched[t];
    if (s.ifsc.count(files[i].id) && nblk[i]) {
      uint8_t *dst = stage + first[i] * bs;
      memset(dst + files[i].length, 0, nblk[i] * bs - files[i].length);
      vreads.push_back({temps[t].first, 0, files[i].length, dst});
      any_ifsc = true;
    }
  }
  if (!io_parallel(vreads, false)) {
    cleanup();
    return fail(ctx, PAR2HIP_EREPAIR, "repair completed but verification failed - files may still be damaged");
  }
  if (any_ifsc) {
    for (size_t t = 0; t < touched.size(); t++) {
      const size_t i = touched[t];
      if (s.ifsc.count(files[i].id) && nblk[i])
        P2H_CHECK(ctx, hipMemcpyAsync((uint8_t *)d_slices + first[i] * bs, stage + first[i] * bs, nblk[i] * bs,
                                      hipMemcpyHostToDevice, ctx->stream));
    }
    if ((rc = par2hip_hash_slices_dev(ctx, d_slices, n, bs, bs, nullptr, d_hash, (uint8_t *)d_hash + n * 16))) {
      cleanup();
      return rc;
    }
    P2H_CHECK(ctx, hipMemcpyAsync(md5s.data(), d_hash, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    P2H_CHECK(ctx, hipMemcpyAsync(crcs.data(), (uint8_t *)d_hash + n * 16, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  for (size_t t = 0; t < touched.size(); t++) {
    const size_t i = touched[t];
    const std::string &tmp = temps[t].first;
    bool ok = true;
    auto it = s.ifsc.find(files[i].id);
    if (it != s.ifsc.end() && nblk[i]) {
      for (uint64_t k = 0; k < nblk[i] && ok; k++) {
        const uint64_t g = first[i] + k;
        ok = k < it->second.size() && crcs[g] == it->second[k].second &&
             memcmp(&md5s[g * 16], it->second[k].first.data(), 16) == 0;
      }
    } else {
      uint8_t d[16];
      ok = md5_file_prefix(tmp, files[i].length, d) && memcmp(d, files[i].md5.data(), 16) == 0;
    }
    if (!ok) {
      cleanup();
      return fail(ctx, PAR2HIP_EREPAIR, "repair completed but verification failed - files may still be damaged");
    }
  }
  // last safe point (repair.rs:1064-1067); once the renames start the
  // repair is committed and the call reports it
  if (cancelled(cancel)) {
    cleanup();
    return fail(ctx, PAR2HIP_ECANCELLED, "cancelled");
  }
  for (size_t t = 0; t < temps.size(); t++) {
    if (rename(temps[t].first.c_str(), temps[t].second.c_str()) != 0) {
      for (size_t u = t; u < temps.size(); u++) (void)remove(temps[u].first.c_str());
      return fail(ctx, PAR2HIP_EINVAL, "cannot rename " + temps[t].first);
    }
  }
  rep->repaired = 1;
  return PAR2HIP_OK;
}

}  // namespace par2hip

extern "C" int32_t par2hip_par2_info(const char *par2_path, par2hip_set_info *info) {
  if (!par2_path || !info) return PAR2HIP_EINVAL;
  return par2hip::set_info(par2_path, info);
}

extern "C" int32_t par2hip_repair_files_cancellable(par2hip_ctx *ctx, const char *par2_path, const char *base_dir,
                                                    int32_t verify_only, const uint8_t *cancel,
                                                    par2hip_repair_report *report) {
  if (!ctx || !par2_path || !report) return PAR2HIP_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  return par2hip::repair_files(ctx, par2_path, base_dir, verify_only, cancel, report);
}

extern "C" int32_t par2hip_repair_files(par2hip_ctx *ctx, const char *par2_path, const char *base_dir,
                                        int32_t verify_only, par2hip_repair_report *report) {
  return par2hip_repair_files_cancellable(ctx, par2_path, base_dir, verify_only, nullptr, report);
}
