// Host-only checks of the PAR2 set loader (par2-rs_amd/csrc/par2set.cpp),
// built by tests/test_par2set_sanitized.py with AddressSanitizer and UBSan:
// the QuickPar fixture, crafted indexes whose sizes would overflow the block
// layout, damaged and truncated copies, garbage, and a bounded random fuzz of
// the fixture's packets.  Usage: par2set_check <fixture dir> <tmp dir> <fuzz iterations>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "md5_host.hpp"
#include "par2set.hpp"

using namespace par2hip;

static int failures = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      failures++;                                                          \
    }                                                                      \
  } while (0)

static std::vector<uint8_t> read_file(const std::string &p) {
  std::vector<uint8_t> b;
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  uint8_t buf[65536];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + k);
  std::fclose(f);
  return b;
}

static void write_file(const std::string &p, const std::vector<uint8_t> &b) {
  FILE *f = std::fopen(p.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", p.c_str());
    std::exit(2);
  }
  if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

static void put64(std::vector<uint8_t> &o, uint64_t v) {
  for (int i = 0; i < 8; i++) o.push_back((uint8_t)(v >> (8 * i)));
}

// One packet with a valid MD5 (PAR 2.0: magic, length, MD5 over set ID,
// type and body, set ID, type, body padded to 4 bytes).
static void packet(std::vector<uint8_t> &o, const uint8_t set_id[16], const char type[16],
                   std::vector<uint8_t> body) {
  body.resize((body.size() + 3) & ~(size_t)3, 0);
  Md5Host h;
  h.update(set_id, 16);
  h.update(type, 16);
  h.update(body.data(), body.size());
  uint8_t d[16];
  h.finish(d);
  const uint8_t magic[8] = {'P', 'A', 'R', '2', 0, 'P', 'K', 'T'};
  o.insert(o.end(), magic, magic + 8);
  put64(o, 64 + body.size());
  o.insert(o.end(), d, d + 16);
  o.insert(o.end(), set_id, set_id + 16);
  o.insert(o.end(), type, type + 16);
  o.insert(o.end(), body.begin(), body.end());
}

static const char kMainT[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'M', 'a', 'i', 'n', 0, 0, 0, 0};
static const char kDescT[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'F', 'i', 'l', 'e', 'D', 'e', 's', 'c'};

// An index with block size bs and one FileDesc per entry of lens.
static std::vector<uint8_t> crafted_index(uint64_t bs, const std::vector<uint64_t> &lens) {
  const uint8_t sid[16] = {7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7};
  std::vector<uint8_t> out, mainb;
  put64(mainb, bs);
  const uint32_t nf = (uint32_t)lens.size();
  for (int i = 0; i < 4; i++) mainb.push_back((uint8_t)(nf >> (8 * i)));
  for (uint32_t f = 0; f < nf; f++)
    for (int i = 0; i < 16; i++) mainb.push_back((uint8_t)(i == 0 ? f + 1 : 0));
  packet(out, sid, kMainT, mainb);
  for (uint32_t f = 0; f < nf; f++) {
    std::vector<uint8_t> d(56, 0);
    d[0] = (uint8_t)(f + 1);  // File ID
    for (int i = 0; i < 8; i++) d[48 + i] = (uint8_t)(lens[f] >> (8 * i));
    const std::string name = "f" + std::to_string(f);
    d.insert(d.end(), name.begin(), name.end());
    packet(out, sid, kDescT, d);
  }
  return out;
}

// Everything the repairer does with a loaded set before touching files.
static void exercise(const std::string &index) {
  Par2Set s;
  std::string err;
  if (load_par2_set(index, s, &err) != PAR2HIP_OK) return;
  const std::vector<FileRec> files = canonical_files(s);
  for (const FileRec &f : files) (void)safe_name(f.name);
  BlockLayout lay;
  if (block_layout(files, s.block_size, 32768, uint64_t(1) << 40, &lay, &err) == PAR2HIP_OK) {
    CHECK(lay.n <= 32768 && lay.bytes == lay.n * s.block_size);
    uint64_t sum = 0;
    for (size_t i = 0; i < files.size(); i++) {
      CHECK(lay.first[i] == sum);
      sum += lay.nblk[i];
    }
    CHECK(sum == lay.n);
  }
  const std::vector<RecSlice> recs = recovery_by_exponent(s);
  for (size_t i = 1; i < recs.size(); i++) CHECK(recs[i - 1].exponent < recs[i].exponent);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <fixture dir> <tmp dir> <fuzz iterations>\n", argv[0]);
    return 2;
  }
  const std::string fx = argv[1], tmp = argv[2];
  const long iters = std::atol(argv[3]);

  // ---- the QuickPar fixture (tests/data of the reference)
  {
    Par2Set s;
    std::string err;
    CHECK(load_par2_set(fx + "/testdata.par2", s, &err) == PAR2HIP_OK);
    CHECK(s.block_size == 5376);
    CHECK(s.files.size() == 10 && s.main_order.size() == 10);
    const std::vector<FileRec> files = canonical_files(s);
    CHECK(files.size() == 10);
    for (size_t i = 0; i < files.size() && i < s.main_order.size(); i++) CHECK(files[i].id == s.main_order[i]);
    BlockLayout lay;
    CHECK(block_layout(files, s.block_size, 32768, uint64_t(1) << 40, &lay, &err) == PAR2HIP_OK);
    CHECK(lay.n == 200);
    const std::vector<RecSlice> recs = recovery_by_exponent(s);
    CHECK(recs.size() == 60 && recs.front().exponent == 0 && recs.back().exponent == 59);
    CHECK(s.ifsc.size() == 10);
  }

  // ---- crafted layouts that overflow 64-bit arithmetic
  {
    std::string err;
    BlockLayout lay;
    std::vector<FileRec> files(4);
    for (auto &f : files) f.length = uint64_t(1) << 62;
    // block size 2^62 (a multiple of 4): 4 blocks whose bytes wrap to 0
    CHECK(block_layout(files, uint64_t(1) << 62, 32768, uint64_t(1) << 40, &lay, &err) == PAR2HIP_EINVAL);
    CHECK(lay.n == 0 && lay.first.empty());
    // block size 4: 2^60 blocks per file
    CHECK(block_layout(files, 4, 32768, uint64_t(1) << 40, &lay, &err) == PAR2HIP_EINVAL);
    // lengths that make the block count wrap
    files.assign(3, FileRec{});
    files[0].length = UINT64_MAX;
    files[1].length = UINT64_MAX;
    CHECK(block_layout(files, 4, UINT64_MAX, UINT64_MAX, &lay, &err) == PAR2HIP_EINVAL);
    CHECK(block_layout(files, 0, 32768, 1, &lay, &err) == PAR2HIP_EINVAL);
    // the largest accepted layout
    files.assign(1, FileRec{});
    files[0].length = 32768ull * 4096;
    CHECK(block_layout(files, 4096, 32768, 32768ull * 4096, &lay, &err) == PAR2HIP_OK && lay.n == 32768);
    CHECK(block_layout(files, 4096, 32767, UINT64_MAX, &lay, &err) == PAR2HIP_EINVAL);
    CHECK(block_layout(files, 4096, 32768, 32768ull * 4096 - 1, &lay, &err) == PAR2HIP_EINVAL);

    // the same through the parser: an index holding block size 2^62 and
    // four 2^62-byte files loads, and its layout is refused
    const std::string p = tmp + "/overflow.par2";
    write_file(p, crafted_index(uint64_t(1) << 62, {1ull << 62, 1ull << 62, 1ull << 62, 1ull << 62}));
    Par2Set s;
    CHECK(load_par2_set(p, s, &err) == PAR2HIP_OK);
    CHECK(s.block_size == (uint64_t(1) << 62) && s.files.size() == 4);
    CHECK(block_layout(canonical_files(s), s.block_size, 32768, uint64_t(1) << 40, &lay, &err) == PAR2HIP_EINVAL);
    write_file(p, crafted_index(4, {1ull << 62}));
    Par2Set s2;
    CHECK(load_par2_set(p, s2, &err) == PAR2HIP_OK);
    CHECK(block_layout(canonical_files(s2), s2.block_size, 32768, uint64_t(1) << 40, &lay, &err) == PAR2HIP_EINVAL);
    // block size not a multiple of 4, and zero
    write_file(p, crafted_index(6, {100}));
    Par2Set s3;
    CHECK(load_par2_set(p, s3, &err) == PAR2HIP_EINVAL);
    write_file(p, crafted_index(0, {100}));
    Par2Set s4;
    CHECK(load_par2_set(p, s4, &err) == PAR2HIP_EINVAL);
    std::remove(p.c_str());
  }

  // ---- stored names
  CHECK(safe_name("a/b.dat") && safe_name("..a") && safe_name("a..") && safe_name("x/..y"));
  CHECK(!safe_name("") && !safe_name("/etc/passwd") && !safe_name("..") && !safe_name("a/../b") &&
        !safe_name("a/.."));

  // ---- block sizes of new sets: the reference's rule (creator.rs:540-585)
  // and the larger size taken where it would pass 32768 blocks
  {
    CHECK(reference_block_size(0) == 2048);
    CHECK(reference_block_size(1000000) == 2048);
    CHECK(reference_block_size(10000000) == 5000);
    CHECK(reference_block_size(1000000000) == 500000);
    const uint64_t huge = 500000000000ull;
    const uint64_t hb = reference_block_size(huge);
    CHECK(hb % 4 == 0 && blocks_of(huge, hb) <= 65535);
    for (unsigned long long t0 : {0ull, 1000ull, 1000000ull, 100000000ull, 10000000000ull, 100ull << 30}) {
      const uint64_t t = t0;
      const uint64_t b = create_block_size(&t, 1, 32768);
      CHECK(b == reference_block_size(t));  // under 128 GiB the rule is the reference's
    }
    for (unsigned long long t0 : {200ull << 30, 129ull << 30, 300ull << 30, (unsigned long long)huge}) {
      const uint64_t t = t0;
      const uint64_t b = create_block_size(&t, 1, 32768);
      CHECK(b % 4 == 0 && b >= reference_block_size(t) && blocks_of(t, b) <= 32768);
      CHECK(blocks_of(t, b - 4) > 32768);  // the smallest that fits
    }
    // many files: each adds its own ceil
    std::vector<uint64_t> lens(1000, (200ull << 30) / 1000 + 1);
    const uint64_t b = create_block_size(lens.data(), lens.size(), 32768);
    uint64_t n = 0, n4 = 0;
    for (uint64_t l : lens) {
      n += blocks_of(l, b);
      n4 += blocks_of(l, b - 4);
    }
    CHECK(b % 4 == 0 && n <= 32768 && n4 > 32768);
    std::vector<uint64_t> many(40000, 1);
    CHECK(create_block_size(many.data(), many.size(), 32768) == 0);
  }

  // ---- damaged copies of the fixture: every packet of the index with one
  // byte flipped, truncations, garbage, a foreign header
  const std::vector<uint8_t> index = read_file(fx + "/testdata.par2");
  const std::vector<uint8_t> vol = read_file(fx + "/testdata.vol00+01.par2");
  CHECK(!index.empty() && !vol.empty());
  const std::string dir = tmp + "/fz";
  std::string cmd = "mkdir -p '" + dir + "'";
  if (std::system(cmd.c_str()) != 0) return 2;
  const std::string ip = dir + "/set.par2", vp = dir + "/set.vol00+01.par2";
  write_file(vp, vol);
  for (size_t off = 0; off < index.size(); off += 97) {
    std::vector<uint8_t> b = index;
    b[off] ^= 0x5A;
    write_file(ip, b);
    exercise(ip);
  }
  for (size_t cut : {size_t(0), size_t(1), size_t(63), size_t(64), size_t(65), index.size() / 2, index.size() - 1}) {
    write_file(ip, std::vector<uint8_t>(index.begin(), index.begin() + (long)std::min(cut, index.size())));
    exercise(ip);
  }
  {
    std::mt19937_64 g(1);
    std::vector<uint8_t> garbage(100000);
    for (auto &x : garbage) x = (uint8_t)g();
    write_file(ip, garbage);
    Par2Set s;
    std::string err;
    CHECK(load_par2_set(ip, s, &err) == PAR2HIP_EINVAL);
  }

  // ---- bounded fuzz: random flips, length-field rewrites, cuts and splices
  std::mt19937_64 rng(12345);
  for (long it = 0; it < iters; it++) {
    std::vector<uint8_t> b = index;
    const int edits = 1 + (int)(rng() % 8);
    for (int e = 0; e < edits && !b.empty(); e++) {
      const size_t at = rng() % b.size();
      switch (rng() % 5) {
        case 0:
          b[at] ^= (uint8_t)(1u << (rng() % 8));
          break;
        case 1: {  // a packet length or a body field set to an extreme value
          static const uint64_t vals[] = {0, 1, 63, 64, 68, UINT64_MAX, UINT64_MAX - 63, 1ull << 62, 1ull << 32};
          const uint64_t v = vals[rng() % 9];
          for (int k = 0; k < 8 && at + k < b.size(); k++) b[at + k] = (uint8_t)(v >> (8 * k));
          break;
        }
        case 2:
          b.resize(at);
          break;
        case 3: {  // splice a piece of the volume in
          const size_t from = rng() % vol.size(), len = std::min<size_t>(rng() % 4096, vol.size() - from);
          b.insert(b.begin() + (long)at, vol.begin() + (long)from, vol.begin() + (long)(from + len));
          break;
        }
        default: {  // duplicate a region
          const size_t len = std::min<size_t>(rng() % 2048, b.size() - at);
          std::vector<uint8_t> piece(b.begin() + (long)at, b.begin() + (long)(at + len));
          b.insert(b.begin() + (long)(rng() % (b.size() + 1)), piece.begin(), piece.end());
        }
      }
    }
    write_file(ip, b);
    if (it % 3 == 0) {  // the volume damaged as well
      std::vector<uint8_t> v = vol;
      v[rng() % v.size()] ^= 0xFF;
      write_file(vp, v);
    }
    exercise(ip);
  }
  std::remove(ip.c_str());
  std::remove(vp.c_str());
  if (failures) {
    std::printf("FAILED %d\n", failures);
    return 1;
  }
  std::printf("OK fuzz=%ld\n", iters);
  return 0;
}
