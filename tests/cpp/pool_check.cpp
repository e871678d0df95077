// Randomized check of the serving loop's best-fit pool against a byte map:
// blocks never overlap, stay inside the range, are granule-aligned; freeing
// everything coalesces back to one block; best fit picks the shortest hole.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "pool.hpp"

using par2hip::kGranule;
using par2hip::RangePool;

static int fail(const char *what, long a = 0, long b = 0) {
  std::printf("FAIL %s %ld %ld\n", what, a, b);
  return 1;
}

int main() {
  const uint64_t size = 1 << 20;  // 4096 granules
  RangePool pool;
  pool.init(size);
  std::vector<int> owner(size / kGranule, -1);
  struct Blk { uint64_t off, len; };
  std::vector<Blk> live;
  std::mt19937_64 rng(7);
  int id = 0;
  for (int step = 0; step < 200000; step++) {
    const bool do_alloc = live.empty() || (rng() % 100) < 55;
    if (do_alloc) {
      const uint64_t want = 1 + rng() % (rng() % 8 == 0 ? 65536 : 4096);
      const uint64_t off = pool.alloc(want);
      const uint64_t len = std::max<uint64_t>((want + kGranule - 1) / kGranule * kGranule, kGranule);
      if (off == UINT64_MAX) {
        if (pool.largest_free() >= len) return fail("alloc refused though a hole fits", (long)want);
        continue;
      }
      if (off % kGranule || off + len > size) return fail("block outside range", (long)off, (long)len);
      for (uint64_t g = off / kGranule; g < (off + len) / kGranule; g++) {
        if (owner[g] != -1) return fail("overlap", (long)off, (long)g);
        owner[g] = id;
      }
      live.push_back({off, len});
      id++;
    } else {
      const size_t k = rng() % live.size();
      const Blk b = live[k];
      if (!pool.release(b.off)) return fail("release of a live block refused", (long)b.off);
      for (uint64_t g = b.off / kGranule; g < (b.off + b.len) / kGranule; g++) owner[g] = -1;
      live[k] = live.back();
      live.pop_back();
    }
    uint64_t used = 0;
    for (const Blk &b : live) used += b.len;
    if (used != pool.in_use()) return fail("in_use mismatch", (long)used, (long)pool.in_use());
  }
  if (pool.release(12345 * kGranule + 1)) return fail("release of a non-block accepted");
  for (const Blk &b : live) pool.release(b.off);
  if (pool.free_blocks() != 1 || pool.largest_free() != size || pool.in_use() != 0)
    return fail("not coalesced", (long)pool.free_blocks(), (long)pool.largest_free());
  // best fit: holes of 3, 1 and 2 granules -> a 1-granule request takes the 1-granule hole
  RangePool p2;
  p2.init(16 * kGranule);
  uint64_t a[8];
  for (int i = 0; i < 8; i++) a[i] = p2.alloc(i == 7 ? 9 * kGranule : kGranule);
  // layout: a0..a6 one granule each, a7 nine granules (fills the rest)
  p2.release(a[0]); p2.release(a[1]); p2.release(a[2]);  // hole of 3 at 0
  p2.release(a[4]);                                       // hole of 1 at 4
  p2.release(a[6]);                                       // hole of 1 at 6
  const uint64_t x = p2.alloc(kGranule);
  if (x != 4 * kGranule) return fail("best fit (lowest offset among shortest)", (long)x);
  std::printf("OK %d allocations\n", id);
  return 0;
}
