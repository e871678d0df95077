// hash_kernels.hip -- batched per-slice IFSC hashes for gfx950.
//
// The reference hashes every input slice, zero-padded to the slice size, with
// MD5 (md-5 0.10.6) and CRC-32/IEEE (crc32fast 1.5.0): src/creator.rs:397-409
// on create, src/verify.rs:548-654 on verify.  Both kernels read slice i at
// data + i*stride; bytes at or beyond lens[i] read as zero (lens == nullptr:
// every slice is full), exactly the reference's zero padding.
//
// MD5 is a serial chain per slice, so one lane owns one slice (64 slices per
// wavefront); its latency per 64-byte block bounds the kernel.  CRC-32 is
// linear, so a large slice is split over a 256-lane workgroup: each lane folds
// a contiguous chunk with slice-by-4 tables held in LDS, and the partial
// remainders are joined by multiplying with x^(8*bytes_after) mod P.  Slices
// under 64 KiB take one lane each (the plain sequential CRC).
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace par2hip {

// ---------------------------------------------------------------- MD5 -----
// RFC 1321 section 3.4 constants: T[i] = floor(2^32 * |sin(i + 1)|).
__constant__ uint32_t kMd5T[64] = {
    0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501,
    0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821,
    0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8,
    0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a,
    0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70,
    0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665,
    0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1,
    0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391};

// Register sets of message words in flight per MD5 lane (prefetch distance
// kMd5Sets-1 blocks of 64 bytes).  Five blocks (~3 us of compression) cover
// the load latency under a concurrent contraction; with two, a 1 GiB encode
// step took 10.13 instead of 9.90 ms.
#ifndef P2H_MD5_SETS
#define P2H_MD5_SETS 6
#endif
constexpr int kMd5Sets = P2H_MD5_SETS;
static_assert(kMd5Sets >= 2, "at least one block must be in flight");

__device__ __forceinline__ uint32_t rotl(uint32_t x, int s) { return __builtin_amdgcn_alignbit(x, x, 32 - s); }

#define MD5_STEP(f, a, b, c, d, m, t, s) \
  a = b + rotl(a + (f) + (m) + (t), s);

__device__ __forceinline__ void md5_compress(uint32_t st[4], const uint32_t M[16]) {
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
  // the four round functions, each one v_bitop3_b32 (truth tables over
  // operands (b, c, d) = (0xF0, 0xCC, 0xAA))
#define F1(b, c, d) __builtin_amdgcn_bitop3_b32((b), (c), (d), 0xCA)  // (b & c) | (~b & d)
#define F2(b, c, d) __builtin_amdgcn_bitop3_b32((b), (c), (d), 0xE4)  // (b & d) | (c & ~d)
#define F3(b, c, d) __builtin_amdgcn_bitop3_b32((b), (c), (d), 0x96)  // b ^ c ^ d
#define F4(b, c, d) __builtin_amdgcn_bitop3_b32((b), (c), (d), 0x39)  // c ^ (b | ~d)
  MD5_STEP(F1(b, c, d), a, b, c, d, M[0], kMd5T[0], 7)
  MD5_STEP(F1(a, b, c), d, a, b, c, M[1], kMd5T[1], 12)
  MD5_STEP(F1(d, a, b), c, d, a, b, M[2], kMd5T[2], 17)
  MD5_STEP(F1(c, d, a), b, c, d, a, M[3], kMd5T[3], 22)
  MD5_STEP(F1(b, c, d), a, b, c, d, M[4], kMd5T[4], 7)
  MD5_STEP(F1(a, b, c), d, a, b, c, M[5], kMd5T[5], 12)
  MD5_STEP(F1(d, a, b), c, d, a, b, M[6], kMd5T[6], 17)
  MD5_STEP(F1(c, d, a), b, c, d, a, M[7], kMd5T[7], 22)
  MD5_STEP(F1(b, c, d), a, b, c, d, M[8], kMd5T[8], 7)
  MD5_STEP(F1(a, b, c), d, a, b, c, M[9], kMd5T[9], 12)
  MD5_STEP(F1(d, a, b), c, d, a, b, M[10], kMd5T[10], 17)
  MD5_STEP(F1(c, d, a), b, c, d, a, M[11], kMd5T[11], 22)
  MD5_STEP(F1(b, c, d), a, b, c, d, M[12], kMd5T[12], 7)
  MD5_STEP(F1(a, b, c), d, a, b, c, M[13], kMd5T[13], 12)
  MD5_STEP(F1(d, a, b), c, d, a, b, M[14], kMd5T[14], 17)
  MD5_STEP(F1(c, d, a), b, c, d, a, M[15], kMd5T[15], 22)

  MD5_STEP(F2(b, c, d), a, b, c, d, M[1], kMd5T[16], 5)
  MD5_STEP(F2(a, b, c), d, a, b, c, M[6], kMd5T[17], 9)
  MD5_STEP(F2(d, a, b), c, d, a, b, M[11], kMd5T[18], 14)
  MD5_STEP(F2(c, d, a), b, c, d, a, M[0], kMd5T[19], 20)
  MD5_STEP(F2(b, c, d), a, b, c, d, M[5], kMd5T[20], 5)
  MD5_STEP(F2(a, b, c), d, a, b, c, M[10], kMd5T[21], 9)
  MD5_STEP(F2(d, a, b), c, d, a, b, M[15], kMd5T[22], 14)
  MD5_STEP(F2(c, d, a), b, c, d, a, M[4], kMd5T[23], 20)
  MD5_STEP(F2(b, c, d), a, b, c, d, M[9], kMd5T[24], 5)
  MD5_STEP(F2(a, b, c), d, a, b, c, M[14], kMd5T[25], 9)
  MD5_STEP(F2(d, a, b), c, d, a, b, M[3], kMd5T[26], 14)
  MD5_STEP(F2(c, d, a), b, c, d, a, M[8], kMd5T[27], 20)
  MD5_STEP(F2(b, c, d), a, b, c, d, M[13], kMd5T[28], 5)
  MD5_STEP(F2(a, b, c), d, a, b, c, M[2], kMd5T[29], 9)
  MD5_STEP(F2(d, a, b), c, d, a, b, M[7], kMd5T[30], 14)
  MD5_STEP(F2(c, d, a), b, c, d, a, M[12], kMd5T[31], 20)

  MD5_STEP(F3(b, c, d), a, b, c, d, M[5], kMd5T[32], 4)
  MD5_STEP(F3(a, b, c), d, a, b, c, M[8], kMd5T[33], 11)
  MD5_STEP(F3(d, a, b), c, d, a, b, M[11], kMd5T[34], 16)
  MD5_STEP(F3(c, d, a), b, c, d, a, M[14], kMd5T[35], 23)
  MD5_STEP(F3(b, c, d), a, b, c, d, M[1], kMd5T[36], 4)
  MD5_STEP(F3(a, b, c), d, a, b, c, M[4], kMd5T[37], 11)
  MD5_STEP(F3(d, a, b), c, d, a, b, M[7], kMd5T[38], 16)
  MD5_STEP(F3(c, d, a), b, c, d, a, M[10], kMd5T[39], 23)
  MD5_STEP(F3(b, c, d), a, b, c, d, M[13], kMd5T[40], 4)
  MD5_STEP(F3(a, b, c), d, a, b, c, M[0], kMd5T[41], 11)
  MD5_STEP(F3(d, a, b), c, d, a, b, M[3], kMd5T[42], 16)
  MD5_STEP(F3(c, d, a), b, c, d, a, M[6], kMd5T[43], 23)
  MD5_STEP(F3(b, c, d), a, b, c, d, M[9], kMd5T[44], 4)
  MD5_STEP(F3(a, b, c), d, a, b, c, M[12], kMd5T[45], 11)
  MD5_STEP(F3(d, a, b), c, d, a, b, M[15], kMd5T[46], 16)
  MD5_STEP(F3(c, d, a), b, c, d, a, M[2], kMd5T[47], 23)

  MD5_STEP(F4(b, c, d), a, b, c, d, M[0], kMd5T[48], 6)
  MD5_STEP(F4(a, b, c), d, a, b, c, M[7], kMd5T[49], 10)
  MD5_STEP(F4(d, a, b), c, d, a, b, M[14], kMd5T[50], 15)
  MD5_STEP(F4(c, d, a), b, c, d, a, M[5], kMd5T[51], 21)
  MD5_STEP(F4(b, c, d), a, b, c, d, M[12], kMd5T[52], 6)
  MD5_STEP(F4(a, b, c), d, a, b, c, M[3], kMd5T[53], 10)
  MD5_STEP(F4(d, a, b), c, d, a, b, M[10], kMd5T[54], 15)
  MD5_STEP(F4(c, d, a), b, c, d, a, M[1], kMd5T[55], 21)
  MD5_STEP(F4(b, c, d), a, b, c, d, M[8], kMd5T[56], 6)
  MD5_STEP(F4(a, b, c), d, a, b, c, M[15], kMd5T[57], 10)
  MD5_STEP(F4(d, a, b), c, d, a, b, M[6], kMd5T[58], 15)
  MD5_STEP(F4(c, d, a), b, c, d, a, M[13], kMd5T[59], 21)
  MD5_STEP(F4(b, c, d), a, b, c, d, M[4], kMd5T[60], 6)
  MD5_STEP(F4(a, b, c), d, a, b, c, M[11], kMd5T[61], 10)
  MD5_STEP(F4(d, a, b), c, d, a, b, M[2], kMd5T[62], 15)
  MD5_STEP(F4(c, d, a), b, c, d, a, M[9], kMd5T[63], 21)
#undef F1
#undef F2
#undef F3
#undef F4
  st[0] += a;
  st[1] += b;
  st[2] += c;
  st[3] += d;
}

// Bytes [pos, pos+64) of a slice whose real length is len, zeros beyond.
__device__ __forceinline__ void md5_load_partial(const P2H_GLOBAL uint8_t *__restrict__ p, uint64_t pos, uint64_t len,
                                                 uint32_t M[16]) {
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t w = 0;
    for (int b = 0; b < 4; b++) {
      const uint64_t o = pos + 4 * k + b;
      if (o < len) w |= (uint32_t)p[o] << (8 * b);
    }
    M[k] = w;
  }
}

// MD5 of one slice: bytes [0, len) of p followed by zeros up to `total`
// (the IFSC zero padding), digest to out[0..3] (nothing stored when out is
// null), continuing from state st after the first `start` bytes (a multiple
// of 64 at or below len).  kAlign: 16 or 4 = p has that byte alignment; 1 =
// unaligned (byte loads).
template <int kAlign>
__device__ __forceinline__ void md5_rest(uint32_t st[4], const P2H_GLOBAL uint8_t *__restrict__ p, uint64_t start,
                                         uint64_t len, uint64_t total, P2H_GLOBAL uint32_t *__restrict__ out) {
  uint32_t M[16];
  const uint64_t full = len & ~(uint64_t)63;  // blocks made entirely of real bytes
  uint64_t pos = start;
  if (kAlign >= 4) {
    // kMd5Sets register sets rotate with a prefetch distance of kMd5Sets-1
    // blocks: the loads of block k+kMd5Sets-1 are issued before block k is
    // compressed.  Every fetch is unconditional (indexes past the end re-read
    // the last block), so each wait covers exactly the set about to be
    // compressed.
    // 32-bit block counts (slices under kMd5MaxSliceBytes = 128 GiB, so
    // every block index and index + kMd5Sets stays below 2^31): the clamp
    // below is then a 32-bit compare, which the scalar unit has for uniform
    // lengths.  Addresses are formed in 64 bits.
    const uint32_t nblk = (uint32_t)((full - start) >> 6);
    p += start;  // block indices below count from start
    if (nblk) {
      const uint32_t last = nblk - 1;
      uint32_t S[kMd5Sets][16];
      auto fetch_at = [&](uint32_t (&D)[16], const P2H_GLOBAL uint8_t *b) {
        if (kAlign == 16) {
          const P2H_GLOBAL u32x4 *r = reinterpret_cast<const P2H_GLOBAL u32x4 *>(b);
          const u32x4 t0 = r[0], t1 = r[1], t2 = r[2], t3 = r[3];
          D[0] = t0.x; D[1] = t0.y; D[2] = t0.z; D[3] = t0.w;
          D[4] = t1.x; D[5] = t1.y; D[6] = t1.z; D[7] = t1.w;
          D[8] = t2.x; D[9] = t2.y; D[10] = t2.z; D[11] = t2.w;
          D[12] = t3.x; D[13] = t3.y; D[14] = t3.z; D[15] = t3.w;
        } else {
          const P2H_GLOBAL uint32_t *r = reinterpret_cast<const P2H_GLOBAL uint32_t *>(b);
#pragma unroll
          for (int k = 0; k < 16; k++) D[k] = r[k];
        }
        // keep the loads here: the scheduler would otherwise sink them to
        // their first use and expose the full memory latency per block
        __builtin_amdgcn_sched_barrier(0);
      };
      auto fetch = [&](uint32_t (&D)[16], uint32_t blk) {
        blk = blk < last ? blk : last;
        fetch_at(D, p + (uint64_t)blk * 64);
      };
#pragma unroll
      for (int j = 0; j < kMd5Sets - 1; j++) fetch(S[j], j);
      uint32_t blk = 0;
      // While every prefetched block lies inside the slice, no clamp is
      // needed: one address per kMd5Sets blocks with immediate offsets serves
      // all the loads, and one wait per block covers its set (loads retire in
      // order, and 4*(kMd5Sets-1) were issued after the set's own four; the
      // 4-byte-aligned form's 16 word loads also compile to four 16-byte
      // loads, and were they not merged the wait would only be stricter).
      // Every instruction here takes one of the chain's issue slots, so this
      // saves about 6 of ~337 per block.
      constexpr int kVmAfter = 4 * (kMd5Sets - 1);
      static_assert(kVmAfter <= 63, "vmcnt holds 6 bits");
      // s_waitcnt immediate: vmcnt in bits 3:0 and 15:14, expcnt (7) and
      // lgkmcnt (15) left at "no wait"
      constexpr int kWaitSet = (kVmAfter & 15) | (7 << 4) | (15 << 8) | ((kVmAfter >> 4) << 14);
      for (; blk + 2 * kMd5Sets - 1 <= nblk; blk += kMd5Sets) {
        const P2H_GLOBAL uint8_t *base = p + (uint64_t)blk * 64;
#pragma unroll
        for (int j = 0; j < kMd5Sets; j++) {
          fetch_at(S[(j + kMd5Sets - 1) % kMd5Sets], base + (j + kMd5Sets - 1) * 64);
          {
            __builtin_amdgcn_s_waitcnt(kWaitSet);
            __builtin_amdgcn_sched_barrier(0);  // the compression stays below the wait
          }
          md5_compress(st, S[j]);
        }
      }
      for (; blk + kMd5Sets <= nblk; blk += kMd5Sets) {
#pragma unroll
        for (int j = 0; j < kMd5Sets; j++) {
          fetch(S[(j + kMd5Sets - 1) % kMd5Sets], blk + j + kMd5Sets - 1);
          md5_compress(st, S[j]);
        }
      }
      // fewer than kMd5Sets blocks remain; they sit in sets 0..kMd5Sets-2
#pragma unroll
      for (int j = 0; j < kMd5Sets - 1; j++)
        if (blk + j < nblk) md5_compress(st, S[j]);
    }
    p -= start;
    pos = full;
  } else {
    for (; pos < full; pos += 64) {
#pragma unroll
      for (int k = 0; k < 16; k++)
        M[k] = (uint32_t)p[pos + 4 * k] | ((uint32_t)p[pos + 4 * k + 1] << 8) |
               ((uint32_t)p[pos + 4 * k + 2] << 16) | ((uint32_t)p[pos + 4 * k + 3] << 24);
      md5_compress(st, M);
    }
  }
  // the block holding the end of the real bytes (if any)
  if (pos < len) {
    md5_load_partial(p, pos, len, M);
    if (pos + 64 <= total) {
      md5_compress(st, M);
      pos += 64;
    }
  }
  // zero blocks of the padding that are followed by at least one more block
  if (pos + 64 <= total) {
#pragma unroll
    for (int k = 0; k < 16; k++) M[k] = 0;
    for (; pos + 64 <= total; pos += 64) md5_compress(st, M);
  }
  // final block(s): remaining real/zero bytes, 0x80, zeros, 64-bit bit length
  {
    const uint32_t rem = (uint32_t)(total - pos);  // < 64
    if (!(pos < len)) {
#pragma unroll
      for (int k = 0; k < 16; k++) M[k] = 0;
    }
    // M already holds bytes [pos, pos+rem) when pos < len (partial load above)
    M[rem >> 2] |= 0x80u << (8 * (rem & 3));
    if (rem >= 56) {
      md5_compress(st, M);
#pragma unroll
      for (int k = 0; k < 16; k++) M[k] = 0;
    }
    const uint64_t bits = total * 8u;
    M[14] = (uint32_t)bits;
    M[15] = (uint32_t)(bits >> 32);
    md5_compress(st, M);
  }
  const u32x4 digest = {st[0], st[1], st[2], st[3]};
  if (out) *reinterpret_cast<P2H_GLOBAL u32x4 *>(out) = digest;
}

template <int kAlign>
__device__ __forceinline__ void md5_slice(const P2H_GLOBAL uint8_t *__restrict__ p, uint64_t len, uint64_t total,
                                          P2H_GLOBAL uint32_t *__restrict__ out) {
  uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
  md5_rest<kAlign>(st, p, 0, len, total, out);
}

// Chains per wave.  A wave with many idle lanes hashed a 1 MiB slice in
// 10.3-11.4 ms against 9.3 ms for a wave of 40 or more chains (1 to 32 active
// lanes, one box, alone on the device), so no wave is left mostly idle: the n
// chains are spread evenly over the ceil(n/64) waves (at least 32 each when
// there are two or more), and the lanes left over recompute one of their own
// wave's chains and store nothing -- every wave runs 64 lanes and no
// wave's longest chain grows.
struct Md5Lane {
  long slice;  // the chain this lane computes
  bool store;  // false: a spare lane's copy
};
__device__ __forceinline__ Md5Lane md5_lane(long n, long lane) {
  const long waves = (n + 63) / 64, w = blockIdx.x;
  const long q = n / waves, rem = n % waves;
  const long cnt = q + (w < rem ? 1 : 0), first = w * q + (w < rem ? w : rem);
  return Md5Lane{first + (lane < cnt ? lane : lane % cnt), lane < cnt};
}

// One lane per slice (64 slices per wavefront).
template <int kAlign>
__global__ __launch_bounds__(64) void k_md5_slices(const uint8_t *__restrict__ data, long n,
                                                   uint64_t slice_size, uint64_t stride,
                                                   const uint64_t *__restrict__ lens,
                                                   uint32_t *__restrict__ out, int high_priority) {
  // when these few chains bound the caller's step, let them issue ahead of
  // co-resident waves of other kernels
  if (high_priority) __builtin_amdgcn_s_setprio(3);
  const Md5Lane ln = md5_lane(n, threadIdx.x);
  const long i = ln.slice;
  const P2H_GLOBAL uint8_t *p = gptr(data) + (uint64_t)i * stride;
  P2H_GLOBAL uint32_t *o = ln.store ? gptr(out) + 4 * i : nullptr;
  if (!lens) {
    // every slice full: the block count and the prefetch clamp are the same
    // for all lanes, so they stay in scalar registers and the loops are
    // uniform (per-lane lengths cost about 4 instructions more per 64-byte
    // block in the clamped loop; the main loop takes the same ~330 either way)
    md5_slice<kAlign>(p, slice_size, slice_size, o);
    return;
  }
  uint64_t len = lens[i];
  if (len > slice_size) len = slice_size;
  md5_slice<kAlign>(p, len, slice_size, o);
}

// Batched form: one lane per entry of a slice list (slices of many sets,
// longest chains first so the waves stay even).
__global__ __launch_bounds__(64) void k_md5_list(const HashEntry *__restrict__ list, long n, int aligned4,
                                                 int high_priority) {
  if (high_priority) __builtin_amdgcn_s_setprio(3);
  const Md5Lane ln = md5_lane(n, threadIdx.x);  // neighbours in the list have similar lengths
  const HashEntry e = list[ln.slice];
  P2H_GLOBAL uint32_t *o = ln.store ? gptr(e.md5) : nullptr;
  if (aligned4)
    md5_slice<4>(gptr(e.data), e.len, e.size, o);
  else
    md5_slice<1>(gptr(e.data), e.len, e.size, o);
}

// ------------------------------------------- MD5 with a helper wave -----
// Each MD5 step adds its message word and its constant, M[g(i)] + T[i],
// which do not depend on the chain.  In this form a second wave of the
// workgroup (the helper) loads the message blocks and writes those 64 sums
// per block and lane to LDS; the chain wave reads them back with 16
// ds_read_b128 per block and issues 4 instructions per step (round function,
// add3, rotate, add) instead of 5: ~280 issue slots per 64-byte block instead
// of ~330, and a wave issues one instruction per 4 clocks, so the chain, the
// step's critical path, shortens by the same fraction.  The sums pass
// through a ring of kMd5Ring chunks of kMd5Chunk blocks (16 KiB per block):
// while the chain reads chunk c (and, near its end, the first block of chunk
// c+1), the helper fills chunk c+2, and one barrier per chunk hands them
// over.  A barrier costs the chain ~120 clocks beyond its issue slot, so the
// chunks are as long as the LDS allows: three blocks (a 144 KiB ring; with
// two, 96 KiB, MD5 of 1024 x 1 MiB alone took 8.19 instead of 8.05 ms).  The
// helper keeps the message words of kMd5HelperSets chunks in flight in
// registers (~3 us of the chain's work).  Whole blocks of full, 16-byte
// aligned slices; the chain hashes the last (< kMd5Chunk) blocks and the
// padding itself.
constexpr int kMd5Chunk = 3, kMd5Ring = 3, kMd5HelperSets = 3;
constexpr int kMd5BlockVecs = 16 * 64;  // u32x4 per block slot: 16 per lane
constexpr int kMd5RingBytes = kMd5Ring * kMd5Chunk * kMd5BlockVecs * 16;

// message word of MD5 step i (RFC 1321 section 3.4: rounds take words k,
// 5k+1, 3k+5 and 7k mod 16)
__device__ constexpr int md5_word(int i) {
  return i < 16 ? i : i < 32 ? (5 * i + 1) & 15 : i < 48 ? (3 * i + 5) & 15 : (7 * i) & 15;
}
__device__ constexpr int md5_rot(int i) {
  return i < 16 ? 7 + 5 * (i & 3) : i < 32 ? (int[4]){5, 9, 14, 20}[i & 3]
       : i < 48 ? (int[4]){4, 11, 16, 23}[i & 3] : (int[4]){6, 10, 15, 21}[i & 3];
}

// md5_compress with MK[i] = M[md5_word(i)] + T[i] given
__device__ __forceinline__ void md5_compress_mk(uint32_t st[4], const uint32_t MK[64]) {
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const uint32_t f = i < 16   ? __builtin_amdgcn_bitop3_b32(b, c, d, 0xCA)
                       : i < 32 ? __builtin_amdgcn_bitop3_b32(b, c, d, 0xE4)
                       : i < 48 ? __builtin_amdgcn_bitop3_b32(b, c, d, 0x96)
                                : __builtin_amdgcn_bitop3_b32(b, c, d, 0x39);
    const uint32_t t = b + rotl(a + f + MK[i], md5_rot(i));
    a = d;
    d = c;
    c = b;
    b = t;
  }
  st[0] += a;
  st[1] += b;
  st[2] += c;
  st[3] += d;
}

// The chunk loop of the helper-wave form, shared by the two kernels below:
// the first nch chunks of the lane's slice p (nch the same for the whole
// workgroup: wave 0 the chains, wave 1 their helper).  The helper returns
// false; the chain returns true with st advanced past those chunks.  kAlign:
// 16 or 4 = p has that byte alignment.
template <int kAlign>
__device__ __forceinline__ bool md5_mk_chunks(const P2H_GLOBAL uint8_t *__restrict__ p, uint32_t nch, int wave,
                                              int lane, uint32_t st[4]) {
  if (nch == 0) return wave == 0;  // uniform: no chunk, no barrier
  __shared__ u32x4 ring[kMd5RingBytes / 16];
  // slot of block blk of a chunk at ring index ri
  auto slot = [&](uint32_t ri, int blk) { return ring + (ri * kMd5Chunk + blk) * kMd5BlockVecs; };

  if (wave == 1) {
    // the helper: message words of chunk ch -> the ring's sums of chunk ch
    uint32_t R[kMd5HelperSets][kMd5Chunk][16];
    auto load = [&](uint32_t (&D)[kMd5Chunk][16], uint32_t ch) {
      ch = ch < nch ? ch : nch - 1;  // unconditional: past the end re-reads the last chunk
      const P2H_GLOBAL uint8_t *b0 = p + (uint64_t)ch * (64 * kMd5Chunk);
      if (kAlign == 16) {
        const P2H_GLOBAL u32x4 *r = reinterpret_cast<const P2H_GLOBAL u32x4 *>(b0);
#pragma unroll
        for (int b = 0; b < kMd5Chunk; b++)
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const u32x4 t = r[4 * b + k];
            D[b][4 * k] = t.x;
            D[b][4 * k + 1] = t.y;
            D[b][4 * k + 2] = t.z;
            D[b][4 * k + 3] = t.w;
          }
      } else {
        const P2H_GLOBAL uint32_t *r = reinterpret_cast<const P2H_GLOBAL uint32_t *>(b0);
#pragma unroll
        for (int b = 0; b < kMd5Chunk; b++)
#pragma unroll
          for (int k = 0; k < 16; k++) D[b][k] = r[16 * b + k];
      }
      __builtin_amdgcn_sched_barrier(0);  // issue the loads here, not at their first use
    };
    auto fill = [&](const uint32_t (&D)[kMd5Chunk][16], uint32_t ch) {
#pragma unroll
      for (int b = 0; b < kMd5Chunk; b++) {
        u32x4 *q = slot(ch % kMd5Ring, b);
#pragma unroll
        for (int v = 0; v < 16; v++) {
          u32x4 t;
          t.x = D[b][md5_word(4 * v)] + kMd5T[4 * v];
          t.y = D[b][md5_word(4 * v + 1)] + kMd5T[4 * v + 1];
          t.z = D[b][md5_word(4 * v + 2)] + kMd5T[4 * v + 2];
          t.w = D[b][md5_word(4 * v + 3)] + kMd5T[4 * v + 3];
          q[v * 64 + lane] = t;
        }
      }
    };
    // chunks 0 and 1 before the first barrier; in step c the helper fills
    // chunk c+2 (a slot the chain reads neither in step c nor before its
    // barrier) and loads chunk c+5 into the registers it just used
    load(R[0], 0);
    load(R[1], 1);
    load(R[2], 2);
    fill(R[0], 0);
    load(R[0], 3);
    fill(R[1], 1);
    load(R[1], 4);
    __syncthreads();
    auto step = [&](uint32_t c, uint32_t (&D)[kMd5Chunk][16]) {
      fill(D, c + 2);
      load(D, c + 5);
      __syncthreads();
    };
    for (uint32_t c = 0; c < nch; c += kMd5HelperSets) {
      step(c, R[2]);
      if (c + 1 < nch) step(c + 1, R[0]);
      if (c + 2 < nch) step(c + 2, R[1]);
    }
    // leave no load in flight when the wave ends
    __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8));
    return false;
  }

  // the chain
  uint32_t S[kMd5Chunk][64];  // one register set per block of a chunk
  auto read = [&](uint32_t (&D)[64], uint32_t ch, int blk) {
    const u32x4 *q = slot(ch, blk);
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const u32x4 t = q[v * 64 + lane];
      D[4 * v] = t.x;
      D[4 * v + 1] = t.y;
      D[4 * v + 2] = t.z;
      D[4 * v + 3] = t.w;
    }
  };
  // Each block's 16 reads go out a whole block before their use and are
  // pinned there (the scheduler would sink them to their uses and expose the
  // LDS latency at every use); one wait before each later read of a chunk,
  // where the previous block's reads are long done, tells the compiler so
  // (s_waitcnt lgkmcnt(0), vmcnt and expcnt left alone), and the barrier's
  // own wait covers the chunk's last read.
  constexpr int kLgkm0 = 15 | (7 << 4) | (0 << 8) | (3 << 14);
  __syncthreads();  // chunks 0 and 1 are in the ring
  read(S[0], 0, 0);
  __builtin_amdgcn_s_waitcnt(kLgkm0);
  __builtin_amdgcn_sched_barrier(0);
  uint32_t ci = 0;  // ring index of chunk c
  for (uint32_t c = 0; c < nch; c++) {
    const uint32_t ni = ci + 1 == kMd5Ring ? 0 : ci + 1;
#pragma unroll
    for (int b = 0; b < kMd5Chunk; b++) {
      if (b > 0) {
        __builtin_amdgcn_s_waitcnt(kLgkm0);
        __builtin_amdgcn_sched_barrier(0);
      }
      // the next chunk was handed over at the previous barrier; after the
      // last chunk this re-reads the current one (a slot the helper no longer
      // writes)
      if (b + 1 < kMd5Chunk) read(S[b + 1], ci, b + 1);
      else read(S[0], c + 1 < nch ? ni : ci, 0);
      __builtin_amdgcn_sched_barrier(0);
      md5_compress_mk(st, S[b]);
    }
    __syncthreads();
    ci = ni;
  }
  return true;
}

// Whole-set form: full slices (slice_size bytes each), 16-byte aligned.
__global__ __launch_bounds__(128) void k_md5_slices_mk(const uint8_t *__restrict__ data, long n,
                                                       uint64_t slice_size, uint64_t stride,
                                                       uint32_t *__restrict__ out, int high_priority) {
  if (high_priority) __builtin_amdgcn_s_setprio(3);
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const Md5Lane ln = md5_lane(n, lane);
  const P2H_GLOBAL uint8_t *p = gptr(data) + (uint64_t)ln.slice * stride;
  // 32-bit block counts: slices stay under kMd5MaxSliceBytes (128 GiB)
  const uint32_t nch = (uint32_t)(slice_size >> 6) / kMd5Chunk;
  uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
  if (!md5_mk_chunks<16>(p, nch, wave, lane, st)) return;
  md5_rest<16>(st, p, (uint64_t)nch * (64 * kMd5Chunk), slice_size, slice_size,
               ln.store ? gptr(out) + 4 * ln.slice : nullptr);
}

// Batched form: one lane per entry of a slice list (4-byte aligned entries,
// longest first so the lanes of a workgroup have similar lengths).  The
// helper covers the whole chunks of real bytes that every lane of the
// workgroup has; each chain hashes the rest of its own entry.
__global__ __launch_bounds__(128) void k_md5_list_mk(const HashEntry *__restrict__ list, long n,
                                                     int high_priority) {
  if (high_priority) __builtin_amdgcn_s_setprio(3);
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const Md5Lane ln = md5_lane(n, lane);
  const HashEntry e = list[ln.slice];
  const P2H_GLOBAL uint8_t *p = gptr(e.data);
  // 32-bit block counts: hashed entries stay under kMd5MaxSliceBytes
  uint32_t nch = (uint32_t)(e.len >> 6) / kMd5Chunk;
#pragma unroll
  for (int o = 32; o; o >>= 1) nch = min(nch, (uint32_t)__shfl_xor((int)nch, o));
  nch = (uint32_t)__builtin_amdgcn_readfirstlane((int)nch);
  uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
  if (!md5_mk_chunks<4>(p, nch, wave, lane, st)) return;
  md5_rest<4>(st, p, (uint64_t)nch * (64 * kMd5Chunk), e.len, e.size, ln.store ? gptr(e.md5) : nullptr);
}

double md5_chains_ms(size_t n, uint64_t bytes, int simds, bool helper) {
  // issue slots per 64-byte block in the chain's main loop, one issue per 4
  // clocks per wave (a wave64 VALU instruction occupies its SIMD 4 clocks):
  // the helper-wave form 281 (256 for the steps, 4 state additions, 16 LDS
  // reads, a wait, loop control and a third of a barrier), the plain form 330 (320
  // for the steps, 4 state additions, 4 loads, a wait, loop control); counted
  // in the compiled loops
  const double blocks = (double)((bytes + 8) / 64 + 1);
  const double waves = (double)((n + 63) / 64);
  const double rounds = std::ceil(waves / std::max(simds, 1));
  return rounds * blocks * (helper ? 281.0 : 330.0) * 4.0 / 2.4e9 * 1e3;
}

bool md5_helper_form(const par2hip_ctx *ctx, hipStream_t s, const void *d_data, size_t n, size_t slice_size,
                     size_t stride, bool lens) {
  if (stride == 0) stride = slice_size;
  // the form holds 144 KiB of LDS, one workgroup per CU: taken for full
  // 16-byte aligned slices of at least one chunk while every chain wave has a
  // CU of its own among the stream's (the hash CUs on the side stream)
  const long stream_cus = (s == ctx->side && ctx->hash_cus > 0) ? ctx->hash_cus : ctx->num_cus;
  return !lens && (uintptr_t)d_data % 16 == 0 && stride % 16 == 0 && slice_size >= 64 * kMd5Chunk &&
         (long)((n + 63) / 64) <= stream_cus;
}

int32_t launch_md5_list(par2hip_ctx *ctx, hipStream_t s, const HashEntry *d_list, size_t n, bool aligned4,
                        bool high_priority) {
  if (n == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "md5_slices", s);
  const long waves = (long)((n + 63) / 64);
  // the helper-wave form while every chain wave has a CU of its own (its
  // ring holds 144 KiB of LDS), as for the whole-set kernel
  const long stream_cus = (s == ctx->side && ctx->hash_cus > 0) ? ctx->hash_cus : ctx->num_cus;
  if (aligned4 && waves <= stream_cus)
    hipLaunchKernelGGL(k_md5_list_mk, dim3((unsigned)waves), dim3(128), 0, s, d_list, (long)n,
                       high_priority ? 1 : 0);
  else
    hipLaunchKernelGGL(k_md5_list, dim3((unsigned)waves), dim3(64), 0, s, d_list, (long)n, aligned4 ? 1 : 0,
                       high_priority ? 1 : 0);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

int32_t launch_md5(par2hip_ctx *ctx, hipStream_t s, const uint8_t *d_data, size_t n,
                   size_t slice_size, size_t stride, const uint64_t *d_lens, uint32_t *d_md5,
                   bool high_priority) {
  if (n == 0) return PAR2HIP_OK;
  if (slice_size >= kMd5MaxSliceBytes) return fail(ctx, PAR2HIP_EINVAL, "slice of 128 GiB or more");
  const bool a16 = ((uintptr_t)d_data % 16 == 0) && (stride % 16 == 0);
  const bool a4 = ((uintptr_t)d_data % 4 == 0) && (stride % 4 == 0);
  ScopedTiming t(ctx, "md5_slices", s);
  dim3 grid((unsigned)((n + 63) / 64));
  if (md5_helper_form(ctx, s, d_data, n, slice_size, stride, d_lens != nullptr))
    hipLaunchKernelGGL(k_md5_slices_mk, grid, dim3(128), 0, s, d_data, (long)n, (uint64_t)slice_size,
                       (uint64_t)stride, d_md5, high_priority ? 1 : 0);
  else if (a16)
    hipLaunchKernelGGL(k_md5_slices<16>, grid, dim3(64), 0, s, d_data, (long)n, (uint64_t)slice_size,
                       (uint64_t)stride, d_lens, d_md5, high_priority ? 1 : 0);
  else if (a4)
    hipLaunchKernelGGL(k_md5_slices<4>, grid, dim3(64), 0, s, d_data, (long)n, (uint64_t)slice_size,
                       (uint64_t)stride, d_lens, d_md5, high_priority ? 1 : 0);
  else
    hipLaunchKernelGGL(k_md5_slices<1>, grid, dim3(64), 0, s, d_data, (long)n, (uint64_t)slice_size,
                       (uint64_t)stride, d_lens, d_md5, high_priority ? 1 : 0);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// ---------------------------------------------------------------- CRC-32 --
// Reflected representation: bit (31-k) holds the coefficient of x^k.
// d_crc_tab: the 256-entry byte table for 0xEDB88320.
__device__ __forceinline__ uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
  uint32_t r = 0;
#pragma unroll 8
  for (int k = 0; k < 32; k++) {
    r ^= (a & (0x80000000u >> k)) ? b : 0u;
    b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
  }
  return r;
}

// One workgroup of 256 lanes per slice.  The (zero-padded) slice is viewed as
// pad leading zero bytes followed by its slice_size bytes, pad chosen so that
// 256 equal chunks of `chunk` bytes (a multiple of 16) cover it; leading
// zeros leave a remainder that starts at zero unchanged.  Each lane folds its
// chunk from a zero register, then a binary tree joins neighbours with
// left * x^(8 * chunk * 2^k) + right (level constants from the host), and the
// initial register 0xFFFFFFFF contributes init_term = ~0 * x^(8*slice_size).



__device__ __forceinline__ uint32_t crc_word(uint32_t crc, uint32_t w, const uint32_t (*T)[256]) {
  crc ^= w;
  return T[3][crc & 0xFF] ^ T[2][(crc >> 8) & 0xFF] ^ T[1][(crc >> 16) & 0xFF] ^ T[0][crc >> 24];
}

// Sequential CRC register update over bytes [o, end) of p, any alignment:
// single bytes up to a 16-byte boundary, then 64-byte blocks of four 16-byte
// loads, then words, then the tail bytes.
__device__ __forceinline__ uint32_t crc_run(uint32_t crc, const P2H_GLOBAL uint8_t *__restrict__ p, uint64_t o, uint64_t end,
                                            const uint32_t (*T)[256]) {
  uint64_t head = (uint64_t)((16 - ((uintptr_t)(p + o) & 15)) & 15);
  if (head > end - o) head = end - o;
  for (const uint64_t h = o + head; o < h; o++) crc = T[0][(crc ^ p[o]) & 0xFF] ^ (crc >> 8);
  for (; o + 64 <= end; o += 64) {
    const P2H_GLOBAL u32x4 *q = reinterpret_cast<const P2H_GLOBAL u32x4 *>(p + o);
    const u32x4 v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3];
    crc = crc_word(crc, v0.x, T); crc = crc_word(crc, v0.y, T); crc = crc_word(crc, v0.z, T); crc = crc_word(crc, v0.w, T);
    crc = crc_word(crc, v1.x, T); crc = crc_word(crc, v1.y, T); crc = crc_word(crc, v1.z, T); crc = crc_word(crc, v1.w, T);
    crc = crc_word(crc, v2.x, T); crc = crc_word(crc, v2.y, T); crc = crc_word(crc, v2.z, T); crc = crc_word(crc, v2.w, T);
    crc = crc_word(crc, v3.x, T); crc = crc_word(crc, v3.y, T); crc = crc_word(crc, v3.z, T); crc = crc_word(crc, v3.w, T);
  }
  for (; o + 4 <= end; o += 4) crc = crc_word(crc, *reinterpret_cast<const P2H_GLOBAL uint32_t *>(p + o), T);
  for (; o < end; o++) crc = T[0][(crc ^ p[o]) & 0xFF] ^ (crc >> 8);
  return crc;
}

// Row layout (the whole kCrcRow rows of a full, 16-byte aligned slice; the
// bytes after them take one chunk per lane): lane l's register runs over its
// kCrcPiece-byte piece of each row,
// the other lanes' bytes counted as zeros.  Skipping the kCrcRow - kCrcPiece
// bytes between two of its pieces multiplies the register by
// x^(8 * (kCrcRow - kCrcPiece)), four lookups in Z (Z_b[v] = (v << 8b) times
// that power).  By linearity the slice's register is the xor over lanes of
// each lane's register times x^(8 * kCrcPiece * (lanes to its right)), which
// the tree join forms with the row levels.  The next row's piece is loaded
// while the current one is folded (two rows ahead measured the same).
__device__ __forceinline__ uint32_t crc_rows(const P2H_GLOBAL uint8_t *__restrict__ p, uint32_t rows,
                                             const uint32_t (*T)[256], const uint32_t (*Z)[256]) {
  const P2H_GLOBAL u32x4 *q = reinterpret_cast<const P2H_GLOBAL u32x4 *>(p + (uint64_t)threadIdx.x * kCrcPiece);
  constexpr uint64_t kRowVecs = kCrcRow / 16;
  u32x4 v[4], w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = q[k];
  uint32_t crc = 0u;
  for (uint32_t r = 0; r < rows; r++) {
    // the next row's piece (the last row reloads its own: in bounds)
    const P2H_GLOBAL u32x4 *nq = q + (uint64_t)(r + 1 < rows ? r + 1 : r) * kRowVecs;
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = nq[k];
    crc = Z[0][crc & 0xFF] ^ Z[1][(crc >> 8) & 0xFF] ^ Z[2][(crc >> 16) & 0xFF] ^ Z[3][crc >> 24];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      crc = crc_word(crc, v[k].x, T);
      crc = crc_word(crc, v[k].y, T);
      crc = crc_word(crc, v[k].z, T);
      crc = crc_word(crc, v[k].w, T);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = w[k];
  }
  return crc;
}

// Register update over n zero bytes.
__device__ __forceinline__ uint32_t crc_zeros(uint32_t crc, uint64_t n, const uint32_t (*T)[256]) {
  for (; n >= 4; n -= 4) crc = crc_word(crc, 0u, T);
  for (; n; n--) crc = T[0][crc & 0xFF] ^ (crc >> 8);
  return crc;
}

// CRC-32 of one slice by the whole 256-lane workgroup (tables T and Z already
// in LDS): bytes [0, len) of p, zeros up to slice_size; result to *out.
__device__ __forceinline__ void crc_slice_wg(const uint32_t (*T)[256], const uint32_t (*Z)[256], uint32_t *part,
                                             const P2H_GLOBAL uint8_t *__restrict__ p, uint64_t len,
                                             uint64_t slice_size, const CrcParams &prm,
                                             P2H_GLOBAL uint32_t *__restrict__ out) {
  const int lane = threadIdx.x;
  // tree join of the lanes' registers: level k merges block [l, l+2^k) with
  // its right neighbour; the total lands in lane 0 (part free again after)
  auto join = [&](uint32_t v, const uint32_t *level) -> uint32_t {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const uint32_t right = __shfl_down(v, 1 << k, 64);
      if ((lane & ((2 << k) - 1)) == 0) v = gf2_mulmod(v, level[k]) ^ right;
    }
    if ((lane & 63) == 0) part[lane >> 6] = v;
    __syncthreads();
    uint32_t tot = 0;
    if (lane == 0) {
      const uint32_t w01 = gf2_mulmod(part[0], level[6]) ^ part[1];
      const uint32_t w23 = gf2_mulmod(part[2], level[6]) ^ part[3];
      tot = gf2_mulmod(w01, level[7]) ^ w23;
    }
    __syncthreads();
    return tot;
  };
  // lane l's bytes of chunk form (chunk, pad) over the n bytes at q, real up
  // to len, zeros after
  auto chunk_reg = [&](const P2H_GLOBAL uint8_t *__restrict__ q, uint64_t chunk, uint64_t pad, uint64_t len) {
    // this lane's bytes: [a, e) real or zero-padding
    const int64_t va = (int64_t)lane * (int64_t)chunk - (int64_t)pad;
    const int64_t ve = va + (int64_t)chunk;
    const uint64_t a = va < 0 ? 0 : (uint64_t)va;
    const uint64_t e = ve < 0 ? 0 : (uint64_t)ve;  // e <= the chunk form's size by construction
    const uint64_t real_end = e < len ? e : len;
    uint32_t c = a < real_end ? crc_run(0u, q, a, real_end, T) : 0u;
    // zero padding inside this lane's range (only when the slice is short)
    const uint64_t zs = a > real_end ? a : real_end;
    if (e > zs) c = crc_zeros(c, e - zs, T);
    return c;
  };
  // the same for the whole workgroup
  if (prm.rows != 0 && len == slice_size && ((uintptr_t)p & 15) == 0) {
    // whole rows, then the tail after them: raw(rows || tail) =
    // raw(rows) x^(8 tail) + raw(tail)
    const uint64_t head = (uint64_t)prm.rows * kCrcRow;
    const uint32_t vr = crc_rows(p, prm.rows, T, Z);
    const uint32_t vt = prm.t_chunk ? chunk_reg(p + head, prm.t_chunk, prm.t_pad, slice_size - head) : 0u;
    const uint32_t tr = join(vr, prm.row_level);
    const uint32_t tt = prm.t_chunk ? join(vt, prm.t_level) : 0u;
    if (lane == 0) *out = ~(prm.init_term ^ (prm.t_chunk ? gf2_mulmod(tr, prm.t_shift) ^ tt : tr));
  } else {
    const uint32_t tot = join(chunk_reg(p, prm.chunk, prm.pad, len), prm.level);
    if (lane == 0) *out = ~(prm.init_term ^ tot);
  }
}

// T_j (j = 0..3, byte followed by j zero bytes) and, with ntab = 8, the row
// skip tables Z_j after them, precomputed on the host: one coalesced copy into
// LDS.
__device__ __forceinline__ void crc_load_tables(uint32_t (*T)[256], const uint32_t *__restrict__ tab, int ntab = 4) {
  for (int k = threadIdx.x; k < 256 * ntab; k += blockDim.x) (&T[0][0])[k] = tab[k];
  __syncthreads();
}

__global__ __launch_bounds__(kCrcLanes) void k_crc32_slices(const uint8_t *__restrict__ data, uint64_t slice_size,
                                                            uint64_t stride, const uint64_t *__restrict__ lens,
                                                            const uint32_t *__restrict__ tab, CrcParams prm,
                                                            uint32_t *__restrict__ out) {
  __shared__ uint32_t T[8][256];  // T0..T3, Z0..Z3
  __shared__ uint32_t part[kCrcLanes / 64];
  crc_load_tables(T, tab, 8);
  const long i = blockIdx.x;
  uint64_t len = lens ? lens[i] : slice_size;
  if (len > slice_size) len = slice_size;
  crc_slice_wg(T, T + 4, part, gptr(data) + (uint64_t)i * stride, len, slice_size, prm, gptr(out) + i);
}

// Batched form: one workgroup per entry; entry e uses params[e.crc_params].
__global__ __launch_bounds__(kCrcLanes) void k_crc32_list(const HashEntry *__restrict__ list,
                                                          const CrcParams *__restrict__ params,
                                                          const uint32_t *__restrict__ tab) {
  __shared__ uint32_t T[8][256];  // T0..T3, Z0..Z3
  __shared__ uint32_t part[kCrcLanes / 64];
  crc_load_tables(T, tab, 8);
  const HashEntry e = list[blockIdx.x];
  const CrcParams prm = params[e.crc_params];
  crc_slice_wg(T, T + 4, part, gptr(e.data), e.len, e.size, prm, gptr(e.crc));
}

static uint32_t host_mulmod(uint32_t a, uint32_t b);
static uint32_t host_x8n(uint64_t nbytes);

int32_t crc_tables_init(par2hip_ctx *ctx) {
  uint32_t h[8][256];
  for (uint32_t k = 0; k < 256; k++) {
    uint32_t c = k;
    for (int b = 0; b < 8; b++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
    h[0][k] = c;
  }
  for (int j = 1; j < 4; j++)
    for (uint32_t k = 0; k < 256; k++) h[j][k] = (h[j - 1][k] >> 8) ^ h[0][h[j - 1][k] & 0xFF];
  // Z_j[v] = (v << 8j) * x^(8 * (kCrcRow - kCrcPiece)): the row layout's skip
  const uint32_t skip = host_x8n(kCrcRow - kCrcPiece);
  for (int j = 0; j < 4; j++)
    for (uint32_t k = 0; k < 256; k++) h[4 + j][k] = host_mulmod(k << (8 * j), skip);
  P2H_CHECK(ctx, hipMalloc(&ctx->d_crc_tab, sizeof h));
  P2H_CHECK(ctx, hipMemcpy(ctx->d_crc_tab, h, sizeof h, hipMemcpyHostToDevice));
  return PAR2HIP_OK;
}

static uint32_t host_mulmod(uint32_t a, uint32_t b) {
  uint32_t r = 0;
  for (int k = 0; k < 32; k++) {
    if (a & (0x80000000u >> k)) r ^= b;
    b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
  }
  return r;
}

static uint32_t host_x8n(uint64_t nbytes) {
  uint32_t r = 0x80000000u, sq = 0x80000000u >> 8;  // 1, x^8
  while (nbytes) {
    if (nbytes & 1) r = host_mulmod(r, sq);
    sq = host_mulmod(sq, sq);
    nbytes >>= 1;
  }
  return r;
}

int32_t launch_crc32(par2hip_ctx *ctx, hipStream_t s, const uint8_t *d_data, size_t n,
                     size_t slice_size, size_t stride, const uint64_t *d_lens, uint32_t *d_crc) {
  if (n == 0) return PAR2HIP_OK;
  const CrcParams prm = crc_params_for(slice_size);
  ScopedTiming t(ctx, "crc32_slices", s);
  hipLaunchKernelGGL(k_crc32_slices, dim3((unsigned)n), dim3(kCrcLanes), 0, s, d_data, (uint64_t)slice_size,
                     (uint64_t)stride, d_lens, ctx->d_crc_tab, prm, d_crc);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

CrcParams crc_params_for(uint64_t slice_size) {
  CrcParams prm;
  prm.chunk = ((slice_size + kCrcLanes - 1) / kCrcLanes + 15) & ~(uint64_t)15;
  if (prm.chunk == 0) prm.chunk = 16;
  prm.pad = prm.chunk * kCrcLanes - slice_size;
  uint32_t lv = host_x8n(prm.chunk);
  for (int k = 0; k < 8; k++) {
    prm.level[k] = lv;
    lv = host_mulmod(lv, lv);
  }
  prm.init_term = host_mulmod(0xFFFFFFFFu, host_x8n(slice_size));
  // row layout over the whole rows (32-bit row count: slices stay under 2^44
  // bytes), the rest in one chunk per lane
  prm.rows = slice_size / kCrcRow < (1ull << 32) ? (uint32_t)(slice_size / kCrcRow) : 0;
  lv = host_x8n(kCrcPiece);
  for (int k = 0; k < 8; k++) {
    prm.row_level[k] = lv;
    lv = host_mulmod(lv, lv);
  }
  const uint64_t tail = slice_size - (uint64_t)prm.rows * kCrcRow;
  prm.t_chunk = tail ? ((tail + kCrcLanes - 1) / kCrcLanes + 15) & ~(uint64_t)15 : 0;
  prm.t_pad = tail ? prm.t_chunk * kCrcLanes - tail : 0;
  lv = host_x8n(prm.t_chunk);
  for (int k = 0; k < 8; k++) {
    prm.t_level[k] = lv;
    lv = host_mulmod(lv, lv);
  }
  prm.t_shift = host_x8n(tail);
  return prm;
}

// Small slices: one lane per entry, the plain sequential CRC (register
// starts at ~0, no combine step).
__global__ __launch_bounds__(256) void k_crc32_list_lane(const HashEntry *__restrict__ list, long n,
                                                        const uint32_t *__restrict__ tab) {
  __shared__ uint32_t T[4][256];
  crc_load_tables(T, tab);
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const HashEntry e = list[i];
  const uint64_t len = e.len < e.size ? e.len : e.size;
  uint32_t crc = crc_run(0xFFFFFFFFu, gptr(e.data), 0, len, T);
  crc = crc_zeros(crc, e.size - len, T);
  *gptr(e.crc) = ~crc;
}

int32_t launch_crc32_list(par2hip_ctx *ctx, hipStream_t s, const HashEntry *d_list, size_t n,
                          const CrcParams *d_params) {
  if (n == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "crc32_slices", s);
  hipLaunchKernelGGL(k_crc32_list, dim3((unsigned)n), dim3(kCrcLanes), 0, s, d_list, d_params, ctx->d_crc_tab);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

int32_t launch_crc32_list_lane(par2hip_ctx *ctx, hipStream_t s, const HashEntry *d_list, size_t n) {
  if (n == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "crc32_small_slices", s);
  hipLaunchKernelGGL(k_crc32_list_lane, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_list, (long)n,
                     ctx->d_crc_tab);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

}  // namespace par2hip
