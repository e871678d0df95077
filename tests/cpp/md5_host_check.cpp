// Prints the host MD5 (md5_host.hpp) of deterministic byte strings, one hex
// digest per line, for comparison with an independent implementation:
//   lengths 0..300 fed in one update, then the same bytes fed in pieces of
//   1, 3, 63, 64 and 65 bytes, then one 5 MiB string.
#include <cstdio>
#include <cstring>
#include <vector>

#include "md5_host.hpp"

static std::vector<uint8_t> bytes(size_t n) {
  std::vector<uint8_t> v(n);
  uint32_t x = 0x12345678u;
  for (size_t i = 0; i < n; i++) {
    x = x * 1664525u + 1013904223u;  // LCG; the checker regenerates the same stream
    v[i] = (uint8_t)(x >> 24);
  }
  return v;
}

static void print(const uint8_t d[16]) {
  for (int i = 0; i < 16; i++) printf("%02x", d[i]);
  printf("\n");
}

int main() {
  const std::vector<uint8_t> all = bytes(5u << 20);
  uint8_t d[16];
  for (size_t n = 0; n <= 300; n++) {
    par2hip::Md5Host::digest(all.data(), n, d);
    print(d);
  }
  const size_t pieces[] = {1, 3, 63, 64, 65};
  for (size_t piece : pieces) {
    par2hip::Md5Host h;
    const size_t n = 1000;
    for (size_t o = 0; o < n; o += piece) h.update(all.data() + o, o + piece <= n ? piece : n - o);
    h.finish(d);
    print(d);
  }
  par2hip::Md5Host::digest(all.data(), all.size(), d);
  print(d);
  return 0;
}
