// par2set.cpp -- host-only PAR2 set loading (see par2set.hpp).
#include "par2set.hpp"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "md5_host.hpp"

namespace par2hip {
namespace {

const uint8_t kMagic[8] = {'P', 'A', 'R', '2', 0, 'P', 'K', 'T'};
const uint8_t kMain[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'M', 'a', 'i', 'n', 0, 0, 0, 0};
const uint8_t kFileDesc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'F', 'i', 'l', 'e', 'D', 'e', 's', 'c'};
const uint8_t kIfsc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'I', 'F', 'S', 'C', 0, 0, 0, 0};
const uint8_t kRecvSlic[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};

uint64_t le64(const uint8_t *p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; i--) v = v << 8 | p[i];
  return v;
}
uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
Id id_at(const uint8_t *p) {
  Id x;
  memcpy(x.data(), p, 16);
  return x;
}

bool regular_file_size(const std::string &p, uint64_t *sz) {
  struct stat st;
  if (stat(p.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return false;
  *sz = (uint64_t)st.st_size;
  return true;
}

// A recovery packet of `len` bytes carries a whole block after its 68-byte
// head (64-byte header + 4-byte exponent); written so that no sum wraps.
bool recovery_packet_fits(uint64_t len, uint64_t block_size) { return len >= 68 && len - 68 >= block_size; }

// The recovery set ID of a .par2 file: that of the packet header at offset
// 0; a file that does not start with one is not a volume of any set
// (parser.rs:606-623).
bool first_set_id(const std::string &path, Id *out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  uint8_t h[64];
  const bool ok = fread(h, 1, 64, f) == 64 && memcmp(h, kMagic, 8) == 0;
  if (ok) *out = id_at(h + 32);
  fclose(f);
  return ok;
}

}  // namespace

std::string dir_of(const std::string &p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? "." : p.substr(0, k);
}

int32_t parse_par2_file(const std::string &path, Par2Set &s, bool recovery_only, std::string *err) {
  uint64_t size = 0;
  if (!regular_file_size(path, &size)) {
    *err = "cannot open " + path;
    return PAR2HIP_EINVAL;
  }
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    *err = "cannot open " + path;
    return PAR2HIP_EINVAL;
  }
  std::vector<std::pair<uint64_t, uint64_t>> deferred;  // recovery packets seen before the block size
  std::vector<uint8_t> body;
  uint64_t pos = 0;
  uint8_t h[64];
  while (size - pos >= 64) {
    if (fseeko(f, (off_t)pos, SEEK_SET) != 0 || fread(h, 1, 64, f) < 64) break;
    if (memcmp(h, kMagic, 8) != 0) {  // resync byte by byte
      pos++;
      continue;
    }
    if (!s.have_set_id && !recovery_only) {
      s.set_id = id_at(h + 32);
      s.have_set_id = true;
    }
    const uint64_t len = le64(h + 8);
    if (len < 64 || len > size - pos) {
      pos++;
      continue;
    }
    // packets are 4-byte aligned; a last packet's padding may run past the end
    const uint64_t padded = len + ((len % 4) ? 4 - len % 4 : 0);
    const uint64_t next = padded > size - pos ? size : pos + padded;
    if (memcmp(h + 48, kRecvSlic, 16) == 0) {
      if (len >= 68) {
        if (s.block_size == 0) {
          deferred.push_back({pos, len});
        } else {
          uint8_t e[4];
          if (fread(e, 1, 4, f) == 4 && recovery_packet_fits(len, s.block_size))
            s.recs.push_back({le32(e), path, pos + 68});
        }
      }
      pos = next;
      continue;
    }
    body.resize(len - 64);
    if (fread(body.data(), 1, body.size(), f) < body.size()) break;
    Md5Host md;
    md.update(h + 32, 32);
    md.update(body.data(), body.size());
    uint8_t digest[16];
    md.finish(digest);
    if (memcmp(digest, h + 16, 16) != 0) {  // a damaged packet is skipped
      pos = next;
      continue;
    }
    if (!recovery_only) {
      if (memcmp(h + 48, kMain, 16) == 0 && body.size() >= 12) {
        const uint64_t bs = le64(body.data());
        const uint32_t nrec = le32(body.data() + 8);
        if ((uint64_t)body.size() >= 12 + 16ull * nrec) {
          s.block_size = bs;
          if (s.main_order.empty())
            for (uint32_t k = 0; k < nrec; k++) s.main_order.push_back(id_at(body.data() + 12 + 16ull * k));
        }
      } else if (memcmp(h + 48, kFileDesc, 16) == 0 && body.size() >= 56) {
        FileRec r;
        r.id = id_at(body.data());
        r.md5 = id_at(body.data() + 16);
        r.md5k = id_at(body.data() + 32);
        r.length = le64(body.data() + 48);
        size_t nl = 0;
        while (56 + nl < body.size() && body[56 + nl]) nl++;
        r.name.assign((const char *)body.data() + 56, nl);
        if (s.files.emplace(r.id, r).second) s.files_in_order.push_back(r);
      } else if (memcmp(h + 48, kIfsc, 16) == 0 && body.size() >= 16) {
        std::vector<std::pair<Id, uint32_t>> ent;
        for (size_t k = 16; k + 20 <= body.size(); k += 20) ent.push_back({id_at(body.data() + k), le32(&body[k + 16])});
        s.ifsc[id_at(body.data())] = ent;
      }
    }
    pos = next;
  }
  for (auto &d : deferred) {  // parser.rs:415-440
    uint8_t e[4];
    if (fseeko(f, (off_t)(d.first + 64), SEEK_SET) == 0 && fread(e, 1, 4, f) == 4 &&
        recovery_packet_fits(d.second, s.block_size))
      s.recs.push_back({le32(e), path, d.first + 68});
  }
  fclose(f);
  return PAR2HIP_OK;
}

int32_t load_par2_set(const std::string &index, Par2Set &s, std::string *err) {
  int32_t rc = parse_par2_file(index, s, false, err);
  if (rc) return rc;
  if (!s.have_set_id) {
    *err = "no packet in " + index;
    return PAR2HIP_EINVAL;
  }
  if (s.block_size == 0 || s.block_size % 4) {  // parser.rs:364-380
    *err = "block size missing, zero or not a multiple of 4";
    return PAR2HIP_EINVAL;
  }
  // the other .par2 files beside the index with the same set ID
  std::vector<std::string> names;
  const std::string dir = dir_of(index);
  const size_t slash = index.find_last_of('/');
  const std::string self = slash == std::string::npos ? index : index.substr(slash + 1);
  if (DIR *dp = opendir(dir.c_str())) {
    while (dirent *e = readdir(dp)) {
      std::string n = e->d_name;
      if (n.size() < 5) continue;
      std::string ext = n.substr(n.size() - 5);
      std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)::tolower(c); });
      if (ext != ".par2" || n == self) continue;
      names.push_back(dir + "/" + n);
    }
    closedir(dp);
  }
  std::sort(names.begin(), names.end());
  for (const std::string &v : names) {
    Id sid;
    std::string ignored;
    if (first_set_id(v, &sid) && sid == s.set_id) (void)parse_par2_file(v, s, true, &ignored);
  }
  return PAR2HIP_OK;
}

std::vector<FileRec> canonical_files(const Par2Set &s) {
  std::vector<FileRec> out;
  std::vector<Id> seen;
  for (const Id &id : s.main_order) {
    auto it = s.files.find(id);
    if (it != s.files.end() && std::find(seen.begin(), seen.end(), id) == seen.end()) {
      out.push_back(it->second);
      seen.push_back(id);
    }
  }
  std::vector<FileRec> rest;
  for (const FileRec &r : s.files_in_order)
    if (std::find(seen.begin(), seen.end(), r.id) == seen.end()) rest.push_back(r);
  std::sort(rest.begin(), rest.end(), [](const FileRec &a, const FileRec &b) {
    for (int i = 15; i >= 0; i--)
      if (a.id[i] != b.id[i]) return a.id[i] < b.id[i];
    return false;
  });
  out.insert(out.end(), rest.begin(), rest.end());
  return out;
}

// The reference joins the stored name to the base path as it stands
// (repair.rs:731, :772), so a crafted set could make it write outside that
// directory; such a set is refused here.
bool safe_name(const std::string &name) {
  if (name.empty() || name[0] == '/') return false;
  size_t b = 0;
  while (b <= name.size()) {
    size_t e = name.find('/', b);
    if (e == std::string::npos) e = name.size();
    if (e - b == 2 && name.compare(b, 2, "..") == 0) return false;
    b = e + 1;
  }
  return true;
}

std::vector<RecSlice> recovery_by_exponent(const Par2Set &s) {
  std::vector<RecSlice> r = s.recs;
  std::stable_sort(r.begin(), r.end(), [](const RecSlice &a, const RecSlice &b) { return a.exponent < b.exponent; });
  r.erase(std::unique(r.begin(), r.end(), [](const RecSlice &a, const RecSlice &b) { return a.exponent == b.exponent; }),
          r.end());
  return r;
}

int32_t block_layout(const std::vector<FileRec> &files, uint64_t block_size, uint64_t max_blocks,
                     uint64_t max_bytes, BlockLayout *out, std::string *err) {
  *out = BlockLayout{};
  if (block_size == 0) {
    *err = "block size is zero";
    return PAR2HIP_EINVAL;
  }
  BlockLayout l;
  l.first.resize(files.size());
  l.nblk.resize(files.size());
  uint64_t n = 0;
  for (size_t i = 0; i < files.size(); i++) {
    const uint64_t k = blocks_of(files[i].length, block_size);
    // n <= max_blocks before the add, so the comparison cannot wrap
    if (k > max_blocks - n) {
      *err = "the set's files need more than " + std::to_string(max_blocks) + " input blocks";
      return PAR2HIP_EINVAL;
    }
    l.first[i] = n;
    l.nblk[i] = k;
    n += k;
  }
  if (n && block_size > max_bytes / n) {
    *err = "the set's blocks (" + std::to_string(n) + " x " + std::to_string(block_size) +
           " bytes) exceed the memory limit of " + std::to_string(max_bytes) + " bytes";
    return PAR2HIP_EINVAL;
  }
  l.n = n;
  l.bytes = n * block_size;
  *out = std::move(l);
  return PAR2HIP_OK;
}

uint64_t reference_block_size(uint64_t total) {
  if (total == 0) return 2048;
  const uint64_t kMin = 2048, kMax = (uint64_t)4 << 20;
  const uint64_t bs = (total / 2000 + 3) & ~(uint64_t)3;
  uint64_t floor_bs = std::max<uint64_t>(blocks_of(total, 65535), kMin);
  floor_bs = (floor_bs + 3) & ~(uint64_t)3;
  if (floor_bs > kMax) return floor_bs;
  return std::max<uint64_t>(std::min<uint64_t>(std::max<uint64_t>(bs, kMin), kMax), floor_bs);
}

uint64_t create_block_size(const uint64_t *file_lens, size_t nfiles, uint64_t max_blocks) {
  uint64_t total = 0, nonempty = 0;
  for (size_t i = 0; i < nfiles; i++) {
    total = file_lens[i] > UINT64_MAX - total ? UINT64_MAX : total + file_lens[i];
    nonempty += file_lens[i] != 0;
  }
  auto count = [&](uint64_t bs) {
    uint64_t n = 0;
    for (size_t i = 0; i < nfiles; i++) n += blocks_of(file_lens[i], bs);
    return n;
  };
  uint64_t bs = reference_block_size(total);
  if (count(bs) <= max_blocks) return bs;
  if (nonempty >= max_blocks) return 0;
  // each non-empty file adds at most one block to total / bs, so this size
  // fits; the steps below only shrink it while the count still fits
  uint64_t hi = (blocks_of(total, max_blocks - nonempty) + 3) & ~(uint64_t)3;
  hi = std::max(hi, bs);
  uint64_t lo = bs;  // count(lo) > max_blocks
  while (hi - lo > 4) {  // smallest multiple of 4 in (lo, hi] that fits
    const uint64_t mid = lo + (((hi - lo) / 2) & ~(uint64_t)3);
    if (count(mid) <= max_blocks) hi = mid;
    else lo = mid;
  }
  return hi;
}

}  // namespace par2hip
