// md5_host.cpp -- the MD5 (RFC 1321) block function of md5_host.hpp, built
// with g++ (the library's host-only sources are): a 1 GiB whole-file digest
// took 1.92 s on the build host from this file against 2.25-2.39 s from the
// same steps in the hipcc host pass (and 2.03 s through the system's
// OpenSSL), and whole-file digests bound PAR2 creation.
#include "md5_host.hpp"

namespace par2hip {

namespace {
inline uint32_t rol(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }
}  // namespace

// The 64 steps of RFC 1321 section 3.4 written out with constant shifts and
// T[i] = floor(2^32 |sin(i+1)|); the round functions in forms with one fewer
// operation: F = d ^ (b & (c ^ d)), G = c ^ (d & (b ^ c)).
void md5_host_blocks(uint32_t st[4], const uint8_t *p, size_t nblocks) {
  for (; nblocks; nblocks--, p += 64) {
    uint32_t m[16];
    memcpy(m, p, 64);  // little-endian words on the host
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
#define P2H_F(x, y, z) ((z) ^ ((x) & ((y) ^ (z))))
#define P2H_G(x, y, z) ((y) ^ ((z) & ((x) ^ (y))))
#define P2H_H(x, y, z) ((x) ^ (y) ^ (z))
#define P2H_I(x, y, z) ((y) ^ ((x) | ~(z)))
#define P2H_STEP(f, w, x, y, z, k, t, s) \
  w += f(x, y, z) + m[k] + (t);          \
  w = rol(w, s) + x;
    // round 2 as (x & z) + (y & ~z): the two terms never share a bit, and the
    // second does not wait for x, the value the previous step just produced
#define P2H_STEP_G(w, x, y, z, k, t, s) \
  w += (y & ~z) + m[k] + (t);           \
  w += x & z;                           \
  w = rol(w, s) + x;
    P2H_STEP(P2H_F, a, b, c, d, 0, 0xd76aa478u, 7)
    P2H_STEP(P2H_F, d, a, b, c, 1, 0xe8c7b756u, 12)
    P2H_STEP(P2H_F, c, d, a, b, 2, 0x242070dbu, 17)
    P2H_STEP(P2H_F, b, c, d, a, 3, 0xc1bdceeeu, 22)
    P2H_STEP(P2H_F, a, b, c, d, 4, 0xf57c0fafu, 7)
    P2H_STEP(P2H_F, d, a, b, c, 5, 0x4787c62au, 12)
    P2H_STEP(P2H_F, c, d, a, b, 6, 0xa8304613u, 17)
    P2H_STEP(P2H_F, b, c, d, a, 7, 0xfd469501u, 22)
    P2H_STEP(P2H_F, a, b, c, d, 8, 0x698098d8u, 7)
    P2H_STEP(P2H_F, d, a, b, c, 9, 0x8b44f7afu, 12)
    P2H_STEP(P2H_F, c, d, a, b, 10, 0xffff5bb1u, 17)
    P2H_STEP(P2H_F, b, c, d, a, 11, 0x895cd7beu, 22)
    P2H_STEP(P2H_F, a, b, c, d, 12, 0x6b901122u, 7)
    P2H_STEP(P2H_F, d, a, b, c, 13, 0xfd987193u, 12)
    P2H_STEP(P2H_F, c, d, a, b, 14, 0xa679438eu, 17)
    P2H_STEP(P2H_F, b, c, d, a, 15, 0x49b40821u, 22)
    P2H_STEP_G(a, b, c, d, 1, 0xf61e2562u, 5)
    P2H_STEP_G(d, a, b, c, 6, 0xc040b340u, 9)
    P2H_STEP_G(c, d, a, b, 11, 0x265e5a51u, 14)
    P2H_STEP_G(b, c, d, a, 0, 0xe9b6c7aau, 20)
    P2H_STEP_G(a, b, c, d, 5, 0xd62f105du, 5)
    P2H_STEP_G(d, a, b, c, 10, 0x02441453u, 9)
    P2H_STEP_G(c, d, a, b, 15, 0xd8a1e681u, 14)
    P2H_STEP_G(b, c, d, a, 4, 0xe7d3fbc8u, 20)
    P2H_STEP_G(a, b, c, d, 9, 0x21e1cde6u, 5)
    P2H_STEP_G(d, a, b, c, 14, 0xc33707d6u, 9)
    P2H_STEP_G(c, d, a, b, 3, 0xf4d50d87u, 14)
    P2H_STEP_G(b, c, d, a, 8, 0x455a14edu, 20)
    P2H_STEP_G(a, b, c, d, 13, 0xa9e3e905u, 5)
    P2H_STEP_G(d, a, b, c, 2, 0xfcefa3f8u, 9)
    P2H_STEP_G(c, d, a, b, 7, 0x676f02d9u, 14)
    P2H_STEP_G(b, c, d, a, 12, 0x8d2a4c8au, 20)
    P2H_STEP(P2H_H, a, b, c, d, 5, 0xfffa3942u, 4)
    P2H_STEP(P2H_H, d, a, b, c, 8, 0x8771f681u, 11)
    P2H_STEP(P2H_H, c, d, a, b, 11, 0x6d9d6122u, 16)
    P2H_STEP(P2H_H, b, c, d, a, 14, 0xfde5380cu, 23)
    P2H_STEP(P2H_H, a, b, c, d, 1, 0xa4beea44u, 4)
    P2H_STEP(P2H_H, d, a, b, c, 4, 0x4bdecfa9u, 11)
    P2H_STEP(P2H_H, c, d, a, b, 7, 0xf6bb4b60u, 16)
    P2H_STEP(P2H_H, b, c, d, a, 10, 0xbebfbc70u, 23)
    P2H_STEP(P2H_H, a, b, c, d, 13, 0x289b7ec6u, 4)
    P2H_STEP(P2H_H, d, a, b, c, 0, 0xeaa127fau, 11)
    P2H_STEP(P2H_H, c, d, a, b, 3, 0xd4ef3085u, 16)
    P2H_STEP(P2H_H, b, c, d, a, 6, 0x04881d05u, 23)
    P2H_STEP(P2H_H, a, b, c, d, 9, 0xd9d4d039u, 4)
    P2H_STEP(P2H_H, d, a, b, c, 12, 0xe6db99e5u, 11)
    P2H_STEP(P2H_H, c, d, a, b, 15, 0x1fa27cf8u, 16)
    P2H_STEP(P2H_H, b, c, d, a, 2, 0xc4ac5665u, 23)
    P2H_STEP(P2H_I, a, b, c, d, 0, 0xf4292244u, 6)
    P2H_STEP(P2H_I, d, a, b, c, 7, 0x432aff97u, 10)
    P2H_STEP(P2H_I, c, d, a, b, 14, 0xab9423a7u, 15)
    P2H_STEP(P2H_I, b, c, d, a, 5, 0xfc93a039u, 21)
    P2H_STEP(P2H_I, a, b, c, d, 12, 0x655b59c3u, 6)
    P2H_STEP(P2H_I, d, a, b, c, 3, 0x8f0ccc92u, 10)
    P2H_STEP(P2H_I, c, d, a, b, 10, 0xffeff47du, 15)
    P2H_STEP(P2H_I, b, c, d, a, 1, 0x85845dd1u, 21)
    P2H_STEP(P2H_I, a, b, c, d, 8, 0x6fa87e4fu, 6)
    P2H_STEP(P2H_I, d, a, b, c, 15, 0xfe2ce6e0u, 10)
    P2H_STEP(P2H_I, c, d, a, b, 6, 0xa3014314u, 15)
    P2H_STEP(P2H_I, b, c, d, a, 13, 0x4e0811a1u, 21)
    P2H_STEP(P2H_I, a, b, c, d, 4, 0xf7537e82u, 6)
    P2H_STEP(P2H_I, d, a, b, c, 11, 0xbd3af235u, 10)
    P2H_STEP(P2H_I, c, d, a, b, 2, 0x2ad7d2bbu, 15)
    P2H_STEP(P2H_I, b, c, d, a, 9, 0xeb86d391u, 21)
#undef P2H_STEP_G
#undef P2H_STEP
#undef P2H_I
#undef P2H_H
#undef P2H_G
#undef P2H_F
    st[0] += a;
    st[1] += b;
    st[2] += c;
    st[3] += d;
  }
}

}  // namespace par2hip
