// md5_host.hpp -- MD5 (RFC 1321) on the host, for the whole-file and packet
// digests of PAR2 creation that are one serial chain each (a CPU core runs a
// chain faster than one GPU lane).  The reference uses md-5 0.10.6
// (Cargo.lock:366-383) at src/creator.rs:80-135 and src/writer.rs:56-110.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace par2hip {

class Md5Host {
 public:
  Md5Host() { reset(); }
  void reset() {
    s_[0] = 0x67452301u;
    s_[1] = 0xefcdab89u;
    s_[2] = 0x98badcfeu;
    s_[3] = 0x10325476u;
    total_ = 0;
    fill_ = 0;
  }
  void update(const void *data, size_t len) {
    const uint8_t *p = static_cast<const uint8_t *>(data);
    total_ += len;
    if (fill_) {
      const size_t take = len < 64 - fill_ ? len : 64 - fill_;
      memcpy(buf_ + fill_, p, take);
      fill_ += take;
      p += take;
      len -= take;
      if (fill_ < 64) return;
      block(buf_);
      fill_ = 0;
    }
    for (; len >= 64; p += 64, len -= 64) block(p);
    memcpy(buf_, p, len);
    fill_ = len;
  }
  void finish(uint8_t out[16]) {
    const uint64_t bits = total_ * 8;
    const uint8_t pad = 0x80;
    update(&pad, 1);
    const uint8_t zero[64] = {0};
    update(zero, (fill_ <= 56) ? 56 - fill_ : 120 - fill_);
    uint8_t len8[8];
    for (int i = 0; i < 8; i++) len8[i] = (uint8_t)(bits >> (8 * i));
    update(len8, 8);
    for (int i = 0; i < 4; i++)
      for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(s_[i] >> (8 * b));
  }
  static void digest(const void *data, size_t len, uint8_t out[16]) {
    Md5Host h;
    h.update(data, len);
    h.finish(out);
  }

 private:
  static uint32_t rol(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }
  void block(const uint8_t *p) {
    // RFC 1321 section 3.4: per-round shifts and T[i] = floor(2^32 |sin(i+1)|)
    static const uint8_t kS[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                                   5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                                   4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                                   6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};
    static const uint32_t kT[64] = {
        0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501,
        0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821,
        0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8,
        0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a,
        0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70,
        0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665,
        0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1,
        0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391};
    uint32_t m[16];
    for (int i = 0; i < 16; i++)
      m[i] = (uint32_t)p[4 * i] | (uint32_t)p[4 * i + 1] << 8 | (uint32_t)p[4 * i + 2] << 16 |
             (uint32_t)p[4 * i + 3] << 24;
    uint32_t a = s_[0], b = s_[1], c = s_[2], d = s_[3];
    for (int i = 0; i < 64; i++) {
      uint32_t f;
      int g;
      if (i < 16) {
        f = (b & c) | (~b & d);
        g = i;
      } else if (i < 32) {
        f = (d & b) | (~d & c);
        g = (5 * i + 1) & 15;
      } else if (i < 48) {
        f = b ^ c ^ d;
        g = (3 * i + 5) & 15;
      } else {
        f = c ^ (b | ~d);
        g = (7 * i) & 15;
      }
      const uint32_t t = d;
      d = c;
      c = b;
      b = b + rol(a + f + kT[i] + m[g], kS[i]);
      a = t;
    }
    s_[0] += a;
    s_[1] += b;
    s_[2] += c;
    s_[3] += d;
  }
  uint32_t s_[4];
  uint64_t total_;
  uint8_t buf_[64];
  size_t fill_;
};

}  // namespace par2hip
