// creator.hip -- PAR2 creation over the GPU entry points: what
// Par2Creator::create (src/creator.rs:261-531) does with the reference's
// encoder and hashers, with the slice hashes, the recovery slices and the
// recovery packets' digests computed on the device.
//
// Host: whole-file and first-16 KiB MD5s (one serial chain per file, threads
// across files), File IDs and the canonical order, the small packets and the
// files.  Device: IFSC MD5/CRC32 of every zero-padded block, the recovery
// slices, and the MD5 of every recovery packet: the contraction writes each
// recovery slice into a packet image row that already holds the set ID, the
// packet type and the exponent, so the packet digest (MD5 over set ID, type
// and body, src/writer.rs:97-105) is one more slice hash.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "md5_host.hpp"
#include "par2set.hpp"

namespace par2hip {
namespace {

const uint8_t kMagic[8] = {'P', 'A', 'R', '2', 0, 'P', 'K', 'T'};
// packet types of the PAR 2.0 specification (src/writer.rs:26-51)
const uint8_t kMain[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'M', 'a', 'i', 'n', 0, 0, 0, 0};
const uint8_t kFileDesc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'F', 'i', 'l', 'e', 'D', 'e', 's', 'c'};
const uint8_t kIfsc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'I', 'F', 'S', 'C', 0, 0, 0, 0};
const uint8_t kRecvSlic[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
const uint8_t kCreator[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'C', 'r', 'e', 'a', 't', 'o', 'r', 0};
// creator text written by the reference (src/creator.rs:496)
const char kCreatorText[] = "par2-rs v0.1.0 https://github.com/zephleggett/par2-rs";

struct Bytes {
  std::vector<uint8_t> v;
  void put(const void *p, size_t n) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    v.insert(v.end(), b, b + n);
  }
  void u32(uint32_t x) {
    for (int i = 0; i < 4; i++) v.push_back((uint8_t)(x >> (8 * i)));
  }
  void u64(uint64_t x) {
    for (int i = 0; i < 8; i++) v.push_back((uint8_t)(x >> (8 * i)));
  }
};

// One packet: header (magic, length, MD5 over set ID + type + body, set ID,
// type) and the body zero-padded to a multiple of 4 (src/writer.rs:83-124).
void append_packet(Bytes &out, const uint8_t type[16], const uint8_t set_id[16], const std::vector<uint8_t> &body) {
  std::vector<uint8_t> padded(body);
  padded.resize((body.size() + 3) & ~(size_t)3, 0);
  Md5Host h;
  h.update(set_id, 16);
  h.update(type, 16);
  h.update(padded.data(), padded.size());
  uint8_t digest[16];
  h.finish(digest);
  out.put(kMagic, 8);
  out.u64(64 + padded.size());
  out.put(digest, 16);
  out.put(set_id, 16);
  out.put(type, 16);
  out.put(padded.data(), padded.size());
}

struct Input {
  const par2hip_input_file *f;
  uint8_t full[16], h16k[16], id[16];
  uint64_t blocks = 0, first = 0;
};

// File ID order of par2cmdline (src/parser.rs:556-564): byte 15 first.
bool id_less(const Input &a, const Input &b) {
  for (int i = 15; i >= 0; i--)
    if (a.id[i] != b.id[i]) return a.id[i] < b.id[i];
  return false;
}

int digits(uint64_t x) {  // src/volumes.rs:47-59
  int w = 1;
  while (x >= 10) {
    x /= 10;
    w++;
  }
  return w;
}

}  // namespace

int32_t create_par2(par2hip_ctx *ctx, const par2hip_input_file *files, size_t nfiles, uint64_t block_size,
                    float redundancy, int32_t scheme, const char *stem, par2hip_output_file *outs, size_t max_outs,
                    size_t *nouts) {
  *nouts = 0;
  if (nfiles == 0 || !stem || !std::isfinite(redundancy) || redundancy < 0.0f ||
      (scheme != PAR2HIP_VOLUMES_SINGLE && scheme != PAR2HIP_VOLUMES_EXPONENTIAL))
    return fail(ctx, PAR2HIP_EINVAL, "create: need files, a stem, a finite redundancy >= 0 and a known scheme");
  for (size_t i = 0; i < nfiles; i++)
    if (!files[i].name || (files[i].len && !files[i].data))
      return fail(ctx, PAR2HIP_EINVAL, "create: every file needs a name and its bytes");
  if (block_size && (block_size < 2048 || block_size % 4))  // with_block_size (creator.rs:185-211)
    return fail(ctx, PAR2HIP_EINVAL, "create: block size must be at least 2048 and a multiple of 4");

  // ---- File IDs (writer.rs:56-63): they need only the first 16 KiB of each
  // file, and fix the block order, the Main packet and the set ID; the
  // whole-file MD5s (creator.rs:80-135, one serial chain each) run on host
  // threads while the device works and are needed only for FileDesc.
  std::vector<Input> in(nfiles);
  for (size_t i = 0; i < nfiles; i++) {
    Input &x = in[i];
    x.f = &files[i];
    Md5Host::digest(x.f->data, std::min<uint64_t>(x.f->len, 16384), x.h16k);
    Md5Host h;
    h.update(x.h16k, 16);
    uint8_t len8[8];
    for (int k = 0; k < 8; k++) len8[k] = (uint8_t)(x.f->len >> (8 * k));
    h.update(len8, 8);
    h.update(x.f->name, strlen(x.f->name));
    h.finish(x.id);
  }
  std::stable_sort(in.begin(), in.end(), id_less);

  // ---- block and recovery counts (creator.rs:291-347)
  if (block_size == 0) {
    std::vector<uint64_t> lens(nfiles);
    for (size_t i = 0; i < nfiles; i++) lens[i] = files[i].len;
    block_size = create_block_size(lens.data(), nfiles, 32768);
    if (block_size == 0) return fail(ctx, PAR2HIP_EINVAL, "create: 32768 or more non-empty files");
  }
  uint64_t n = 0;
  for (Input &x : in) {
    x.first = n;
    x.blocks = (x.f->len + block_size - 1) / block_size;
    n += x.blocks;
  }
  const uint64_t recovery_bytes = (uint64_t)((double)(n * block_size) * ((double)redundancy / 100.0));
  const uint64_t r = (recovery_bytes + block_size - 1) / block_size;
  if (n > 32768) return fail(ctx, PAR2HIP_EINVAL, "create: more than 32768 input blocks repeat PAR2 constants");
  if (r > 65535 || n + r > 65535) return fail(ctx, PAR2HIP_EINVAL, "create: data + recovery blocks exceed 65535");
  if (scheme == PAR2HIP_VOLUMES_SINGLE && r == 0)  // creator.rs:455-457 needs one volume
    return fail(ctx, PAR2HIP_EINVAL, "create: the single-file scheme needs at least one recovery block");

  // ---- output files and their exponent ranges (volumes.rs:90-208, creator.rs:453-469)
  struct Vol {
    std::string path;
    uint64_t start, end;  // [start, end)
  };
  std::vector<Vol> vols;
  vols.push_back({std::string(stem) + ".par2", 0, scheme == PAR2HIP_VOLUMES_SINGLE ? r : 0});
  if (scheme == PAR2HIP_VOLUMES_EXPONENTIAL) {
    const int w = digits(r ? r - 1 : 0);
    for (uint64_t e = 0, want = 1; e < r; want = std::min<uint64_t>(want * 2, 1024)) {
      const uint64_t cnt = std::min(want, r - e);
      char name[64];
      snprintf(name, sizeof name, ".vol%0*llu+%0*llu.par2", w, (unsigned long long)e, w, (unsigned long long)cnt);
      vols.push_back({std::string(stem) + name, e, e + cnt});
      e += cnt;
    }
  }
  *nouts = vols.size();
  if (vols.size() > max_outs) return fail(ctx, PAR2HIP_EINVAL, "create: output array too small");

  // whole-file digests, in the background (threads across files)
  std::vector<std::thread> hashers;
  {
    const size_t nt = std::max<size_t>(1, std::min<size_t>(nfiles, std::thread::hardware_concurrency()));
    for (size_t t = 0; t < nt; t++)
      hashers.emplace_back([&in, nt, t] {
        for (size_t i = t; i < in.size(); i += nt) Md5Host::digest(in[i].f->data, in[i].f->len, in[i].full);
      });
  }
  struct Joiner {
    std::vector<std::thread> &th;
    ~Joiner() {
      for (auto &t : th)
        if (t.joinable()) t.join();
    }
  } joiner{hashers};

  // ---- Main packet body and recovery set ID (writer.rs:137-173, :66-71)
  Bytes mainb;
  mainb.u64(block_size);
  mainb.u32((uint32_t)nfiles);
  for (const Input &x : in) mainb.put(x.id, 16);
  uint8_t set_id[16];
  Md5Host::digest(mainb.v.data(), mainb.v.size(), set_id);

  // ---- device: blocks in canonical order, IFSC hashes, recovery packet image.
  // The blocks go through the context's pinned staging buffer (filled by
  // several threads, one DMA); the packet image comes back through it.
  const uint64_t row = 36 + block_size;  // set ID, type, exponent, slice
  std::vector<uint8_t> md5s(n * 16), pkt_md5(r * 16), host_img;
  std::vector<uint32_t> crcs(n);
  const uint8_t *img = nullptr;  // r rows of `row` bytes
  int32_t rc;
  if (n) {
    uint8_t *stage = nullptr;
    void *d_blocks, *d_hash, *d_img = nullptr, *d_pkt = nullptr;
    if ((rc = host_stage(ctx, std::max(n * block_size, r * row), &stage))) return rc;
    if ((rc = scratch(ctx, kScratchIn, n * block_size, &d_blocks))) return rc;
    if ((rc = scratch(ctx, kScratchHash, n * 20, &d_hash))) return rc;
    if (r && ((rc = scratch(ctx, kScratchAux, r * row, &d_img)) || (rc = scratch(ctx, kScratchAux2, r * 16, &d_pkt))))
      return rc;
    std::vector<CopyJob> copies;
    for (const Input &x : in) {
      uint8_t *dst = stage + x.first * block_size;
      if (x.f->len) copies.push_back({dst, static_cast<const uint8_t *>(x.f->data), x.f->len});
      memset(dst + x.f->len, 0, x.blocks * block_size - x.f->len);  // zero padding of the last block
    }
    copy_parallel(copies, 8);
    P2H_CHECK(ctx, hipMemcpyAsync(d_blocks, stage, n * block_size, hipMemcpyHostToDevice, ctx->stream));
    if (r) {
      std::vector<uint8_t> prefix(r * 36);
      for (uint64_t e = 0; e < r; e++) {
        memcpy(&prefix[e * 36], set_id, 16);
        memcpy(&prefix[e * 36 + 16], kRecvSlic, 16);
        for (int k = 0; k < 4; k++) prefix[e * 36 + 32 + k] = (uint8_t)(e >> (8 * k));
      }
      P2H_CHECK(ctx, hipMemcpy2DAsync(d_img, row, prefix.data(), 36, 36, r, hipMemcpyHostToDevice, ctx->stream));
      // the prefix buffer must stay alive until the copy has run
      P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    std::vector<uint32_t> exps(r);
    for (uint64_t e = 0; e < r; e++) exps[e] = (uint32_t)e;
    if ((rc = encode_dev_impl(ctx, (const uint8_t *)d_blocks, n, block_size, exps.data(), r,
                              r ? (uint8_t *)d_img + 36 : nullptr, (uint32_t *)d_hash,
                              (uint32_t *)((uint8_t *)d_hash + n * 16), row)))
      return rc;
    if (r && (rc = launch_md5(ctx, ctx->stream, (const uint8_t *)d_img, r, row, row, nullptr, (uint32_t *)d_pkt,
                              false)))
      return rc;
    P2H_CHECK(ctx, hipMemcpyAsync(md5s.data(), d_hash, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    P2H_CHECK(ctx, hipMemcpyAsync(crcs.data(), (uint8_t *)d_hash + n * 16, n * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
    if (r) {
      // the staging buffer's blocks have reached the device (same stream)
      P2H_CHECK(ctx, hipMemcpyAsync(stage, d_img, r * row, hipMemcpyDeviceToHost, ctx->stream));
      P2H_CHECK(ctx, hipMemcpyAsync(pkt_md5.data(), d_pkt, r * 16, hipMemcpyDeviceToHost, ctx->stream));
    }
    P2H_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    img = stage;
  } else if (r) {
    // no data blocks: every recovery slice is zero
    host_img.assign(r * row, 0);
    for (uint64_t e = 0; e < r; e++) {
      uint8_t *p = &host_img[e * row];
      memcpy(p, set_id, 16);
      memcpy(p + 16, kRecvSlic, 16);
      for (int k = 0; k < 4; k++) p[32 + k] = (uint8_t)(e >> (8 * k));
      Md5Host::digest(p, row, &pkt_md5[e * 16]);
    }
    img = host_img.data();
  }

  // ---- the packets every file starts with (creator.rs:483-516).  Only the
  // FileDesc packets need the whole-file MD5s; the head's size does not
  // depend on them, so each file's recovery packets are written at their
  // final offset while the MD5 threads finish, and the head last.
  auto build_head = [&](bool with_full) {
    Bytes head;
    append_packet(head, kCreator, set_id, std::vector<uint8_t>(kCreatorText, kCreatorText + strlen(kCreatorText)));
    append_packet(head, kMain, set_id, mainb.v);
    const uint8_t zero[16] = {};
    for (const Input &x : in) {
      Bytes b;
      b.put(x.id, 16);
      b.put(with_full ? x.full : zero, 16);
      b.put(x.h16k, 16);
      b.u64(x.f->len);
      b.put(x.f->name, strlen(x.f->name) + 1);
      append_packet(head, kFileDesc, set_id, b.v);
    }
    for (const Input &x : in) {
      Bytes b;
      b.put(x.id, 16);
      for (uint64_t k = 0; k < x.blocks; k++) {
        b.put(&md5s[(x.first + k) * 16], 16);
        b.u32(crcs[x.first + k]);
      }
      append_packet(head, kIfsc, set_id, b.v);
    }
    return head;
  };
  const uint64_t head_size = build_head(false).v.size();

  // ---- the files: one writer thread per file for its recovery packets
  std::vector<int> fds(vols.size(), -1);
  struct Closer {
    std::vector<int> &fds;
    ~Closer() {
      for (int fd : fds)
        if (fd >= 0) (void)close(fd);
    }
  } closer{fds};
  for (size_t i = 0; i < vols.size(); i++) {
    fds[i] = open(vols[i].path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0666);
    if (fds[i] < 0) return fail(ctx, PAR2HIP_EINVAL, "create: cannot write " + vols[i].path);
  }
  auto pwrite_all = [](int fd, const uint8_t *p, uint64_t len, uint64_t off) {
    while (len) {
      const ssize_t w = pwrite(fd, p, len, (off_t)off);
      if (w <= 0) return false;
      p += w;
      off += (uint64_t)w;
      len -= (uint64_t)w;
    }
    return true;
  };
  std::vector<char> wrote(vols.size(), 1);
  std::vector<std::thread> writers;
  for (size_t i = 0; i < vols.size(); i++)
    writers.emplace_back([&, i] {
      const Vol &v = vols[i];
      uint64_t off = head_size;
      for (uint64_t e = v.start; e < v.end && wrote[i]; e++) {
        Bytes hdr;
        hdr.put(kMagic, 8);
        hdr.u64(64 + 4 + block_size);
        hdr.put(&pkt_md5[e * 16], 16);
        wrote[i] = pwrite_all(fds[i], hdr.v.data(), hdr.v.size(), off) &&
                   pwrite_all(fds[i], img + e * row, row, off + hdr.v.size());
        off += hdr.v.size() + row;
      }
    });
  for (auto &t : hashers) t.join();
  const Bytes head = build_head(true);
  bool ok = head.v.size() == head_size;
  for (size_t i = 0; i < vols.size() && ok; i++) ok = pwrite_all(fds[i], head.v.data(), head.v.size(), 0);
  for (auto &t : writers) t.join();
  // writeback started and files closed in parallel: ext4 forces the block
  // allocation of a file truncated and rewritten at its close, one file after
  // another if the closes run on one thread
  std::vector<char> closed(vols.size(), 0);
  std::vector<std::thread> closers;
  for (size_t i = 0; i < vols.size(); i++)
    closers.emplace_back([&, i] {
      (void)sync_file_range(fds[i], 0, 0, SYNC_FILE_RANGE_WRITE);
      closed[i] = close(fds[i]) == 0;
    });
  for (auto &t : closers) t.join();
  for (size_t i = 0; i < vols.size(); i++) {
    ok = ok && wrote[i] && closed[i];
    fds[i] = -1;
  }
  if (!ok) return fail(ctx, PAR2HIP_EINVAL, "create: write failed for " + vols[0].path + " or a volume");
  for (size_t i = 0; i < vols.size(); i++) {
    snprintf(outs[i].path, sizeof outs[i].path, "%s", vols[i].path.c_str());
    outs[i].bytes = head_size + (vols[i].end - vols[i].start) * (64 + 4 + block_size);
  }
  return PAR2HIP_OK;
}

}  // namespace par2hip

extern "C" int32_t par2hip_create_par2(par2hip_ctx *ctx, const par2hip_input_file *files, size_t nfiles,
                                       uint64_t block_size, float redundancy_percent, int32_t scheme,
                                       const char *out_stem, par2hip_output_file *outs, size_t max_outs,
                                       size_t *nouts) {
  if (!ctx || !nouts || (max_outs && !outs) || (nfiles && !files)) return PAR2HIP_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  return par2hip::create_par2(ctx, files, nfiles, block_size, redundancy_percent, scheme, out_stem, outs, max_outs,
                              nouts);
}
