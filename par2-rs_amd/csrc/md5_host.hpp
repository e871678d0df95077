// md5_host.hpp -- MD5 (RFC 1321) on the host, for the whole-file and packet
// digests of PAR2 creation that are one serial chain each (a CPU core runs a
// chain faster than one GPU lane).  The reference uses md-5 0.10.6
// (Cargo.lock:366-383) at src/creator.rs:80-135 and src/writer.rs:56-110.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace par2hip {

// nblocks consecutive 64-byte blocks of p into the MD5 state st (A, B, C, D).
// Out of line in md5_host.cpp, which the library builds with g++: the
// hipcc host pass compiled the same steps 12-17% slower.
void md5_host_blocks(uint32_t st[4], const uint8_t *p, size_t nblocks);

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
    if (len >= 64) {
      md5_host_blocks(s_, p, len / 64);
      p += len & ~(size_t)63;
      len &= 63;
    }
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
  void block(const uint8_t *p) { md5_host_blocks(s_, p, 1); }
  uint32_t s_[4];
  uint64_t total_;
  uint8_t buf_[64];
  size_t fill_;
};

}  // namespace par2hip
