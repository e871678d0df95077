// gf_device.hpp -- device code shared by the single-set and batched
// contraction kernels (variant T, LDS nibble tables).
#pragma once
#include "common.hpp"

namespace par2hip {

constexpr int kTpb = 256;  // threads per workgroup

// Tile shape of the contraction: V words per thread, G groups of 8 output
// rows per workgroup (accumulators: G*V 16-byte registers), D sources per
// LDS table chunk (D*G KiB of LDS).

// c * x and c * x^4 in GF(2^16) mod 0x1100B (x^16 = x^12 + x^3 + x + 1, the
// low 16 bits of the polynomial): the bits shifted out at the top come back
// multiplied by 0x100B, which for at most 4 bits needs no further reduction.
__device__ __forceinline__ uint32_t gf_mul_x(uint32_t c) {
  return ((c << 1) & 0xFFFFu) ^ ((c >> 15) ? (kPoly & 0xFFFFu) : 0u);
}
__device__ __forceinline__ uint32_t gf_mul_x4(uint32_t c) {
  const uint32_t t = c >> 12;
  return ((c << 4) & 0xFFFFu) ^ (t << 12) ^ (t << 3) ^ (t << 1) ^ t;
}

// The 64 table entries of one (8-row group, source) pair, written by one
// wavefront without table lookups.  Lane (j, i, h) = (row j, nibble
// position i, high bit h of the nibble value) forms c_j * x^(4i), its
// multiples by x, x^2, x^3, and from those the products for the nibble
// values v = 8h..8h+7 in Gray-code order (one XOR each); it stores them as
// row j of entries 16i + v.  Rows past ndst get zero products.
__device__ __forceinline__ void gf_tab_pair(P2H_GLOBAL uint16_t *__restrict__ tab, const uint16_t *__restrict__ coeff,
                                            int ldc, int ndst, int g, int d, int lane) {
  const int j = lane >> 3, i = (lane >> 1) & 3, h = lane & 1;
  const int r = kTabRows * g + j;
  uint32_t c = r < ndst ? coeff[(long)r * ldc + d] : 0u;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint32_t t = gf_mul_x4(c);
    c = k < i ? t : c;
  }
  const uint32_t b0 = c, b1 = gf_mul_x(b0), b2 = gf_mul_x(b1), b3 = gf_mul_x(b2);
  // u16 index of row j in entry 16i + 8h + u is 8 * (16i + 8h + u) + j
  P2H_GLOBAL uint16_t *q = tab + 8 * (16 * i + 8 * h) + j;
  uint32_t v = h ? b3 : 0u;            // u = 0
  q[8 * 0] = (uint16_t)v;
  v ^= b0; q[8 * 1] = (uint16_t)v;     // u = 1
  v ^= b1; q[8 * 3] = (uint16_t)v;     // u = 3
  v ^= b0; q[8 * 2] = (uint16_t)v;     // u = 2
  v ^= b2; q[8 * 6] = (uint16_t)v;     // u = 6
  v ^= b0; q[8 * 7] = (uint16_t)v;     // u = 7
  v ^= b1; q[8 * 5] = (uint16_t)v;     // u = 5
  v ^= b0; q[8 * 4] = (uint16_t)v;     // u = 4
}

template <int V>
struct WordVec;
template <>
struct WordVec<8> { using T = u32x4; };
template <>
struct WordVec<4> { using T = u32x2; };
template <>
struct WordVec<2> { using T = uint32_t; };
template <>
struct WordVec<1> { using T = uint16_t; };

// How a thread moves its V words: 0 = one vector access (the whole range is
// inside the row and aligned to 2V bytes), 1 = dword pairs (inside the row,
// 4-byte aligned), 2 = word by word with bounds checks (the last, partial
// tile; rows aligned to 2 bytes only).  Modes 0 and 1 issue unconditional
// accesses, so the wait before a use covers exactly the load it needs.
// Words stay packed two per dword (little-endian: word 2i in the low half of
// r[i]) until they are used.
template <int V, int kMode>
__device__ __forceinline__ void load_words(const P2H_GLOBAL uint16_t *__restrict__ p, long w0, long words,
                                           uint32_t (&r)[(V + 1) / 2]) {
  using Vec = typename WordVec<V>::T;
  if constexpr (kMode == 0) {
    const Vec t = *reinterpret_cast<const P2H_GLOBAL Vec *>(p + w0);
    if constexpr (V == 1)
      r[0] = t;
    else
      __builtin_memcpy(r, &t, sizeof t);
  } else if constexpr (kMode == 1) {
#pragma unroll
    for (int k = 0; k < V; k += 2) r[k / 2] = *reinterpret_cast<const P2H_GLOBAL uint32_t *>(p + w0 + k);
  } else {
#pragma unroll
    for (int k = 0; k < (V + 1) / 2; k++) r[k] = 0;
#pragma unroll
    for (int k = 0; k < V; k++)
      if (w0 + k < words) r[k / 2] |= (uint32_t)p[w0 + k] << (16 * (k & 1));
  }
}

template <int V, int kMode>
__device__ __forceinline__ void store_words(P2H_GLOBAL uint16_t *__restrict__ p, long w0, long words,
                                            const uint32_t (&x)[V], bool accumulate) {
  using Vec = typename WordVec<V>::T;
  if constexpr (kMode == 0) {
    Vec t;
    uint16_t *h = reinterpret_cast<uint16_t *>(&t);
    if (accumulate) {
      t = *reinterpret_cast<const P2H_GLOBAL Vec *>(p + w0);
#pragma unroll
      for (int k = 0; k < V; k++) h[k] ^= (uint16_t)x[k];
    } else {
#pragma unroll
      for (int k = 0; k < V; k++) h[k] = (uint16_t)x[k];
    }
    *reinterpret_cast<P2H_GLOBAL Vec *>(p + w0) = t;
  } else if constexpr (kMode == 1) {
#pragma unroll
    for (int k = 0; k < V; k += 2) {
      P2H_GLOBAL uint32_t *q = reinterpret_cast<P2H_GLOBAL uint32_t *>(p + w0 + k);
      uint32_t t = (x[k] & 0xFFFFu) | (x[k + 1] << 16);
      if (accumulate) t ^= *q;
      *q = t;
    }
  } else {
#pragma unroll
    for (int k = 0; k < V; k++)
      if (w0 + k < words) p[w0 + k] = accumulate ? (uint16_t)(p[w0 + k] ^ x[k]) : (uint16_t)x[k];
  }
}

// acc ^= (tables of source d in lds) x (its V words), the words packed two per
// dword.  One 16-byte LDS read per nibble and group (8 rows); the four
// partial products and the accumulator are folded by two three-input XORs.
template <int V, int G>
__device__ __forceinline__ void gf_lds_lookup(u32x4 (&acc)[G][V], const u32x4 *__restrict__ t,
                                              const uint32_t (&r)[(V + 1) / 2]) {
#pragma unroll
  for (int k = 0; k < V; k++) {
    const uint32_t x = r[k / 2] >> (16 * (k & 1));
    const uint32_t a0 = x & 15u, a1 = 16u + ((x >> 4) & 15u);
    const uint32_t a2 = 32u + ((x >> 8) & 15u), a3 = 48u + ((x >> 12) & 15u);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const u32x4 *tg = t + g * 64;
      const u32x4 l0 = tg[a0], l1 = tg[a1], l2 = tg[a2], l3 = tg[a3];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t h = __builtin_amdgcn_bitop3_b32(acc[g][k][q], l0[q], l1[q], 0x96);
        acc[g][k][q] = __builtin_amdgcn_bitop3_b32(h, l2[q], l3[q], 0x96);
      }
    }
  }
}

// acc ^= sum over sources d0..d0+dn-1 of (their tables in lds) x (their
// words).  Memory latency is hidden by the other waves of the CU.
template <int V, int G, int kMode>
__device__ __forceinline__ void gf_lds_sources(u32x4 (&acc)[G][V], const u32x4 *__restrict__ lds,
                                               const uint16_t *const *__restrict__ srcs, int d0, int dn, long w0,
                                               long words) {
  const auto sp = gptr(srcs) + d0;  // the row pointers, read from global memory
  for (int d = 0; d < dn; d++) {
    uint32_t r[(V + 1) / 2];
    load_words<V, kMode>(gptr(sp[d]), w0, words, r);
    gf_lds_lookup<V, G>(acc, lds + d * G * 64, r);
  }
}

// The same with the words of source d+1 in flight while the tables of
// source d are looked up (the last load repeats the last row).  The stage's
// row pointers arrive one per lane (lane k: source d0+k, loaded by the
// caller while every lane is active) and are read back with readlane, so the
// loop issues no scalar loads: one in flight returns out of order and turns
// every wait on the table reads into a wait for all of them.  It costs ~20
// VGPRs (one wave per SIMD less at the default shape).  Measured on one
// device: 5% faster for the persistent batch kernel and 9% for a 1024-source
// launch, 6% slower for a 2000-source launch of the default shape, so only
// the batch kernel takes it.
template <int V, int G, int kMode>
__device__ __forceinline__ void gf_lds_sources_pf(u32x4 (&acc)[G][V], const u32x4 *__restrict__ lds, uint32_t plo,
                                                  uint32_t phi, int dn, long w0, long words) {
  auto row = [&](int d) {
    const uint64_t a = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)phi, d) << 32 |
                       (uint32_t)__builtin_amdgcn_readlane((int)plo, d);
    return reinterpret_cast<const P2H_GLOBAL uint16_t *>(a);
  };
  uint32_t r[(V + 1) / 2], rn[(V + 1) / 2];
  load_words<V, kMode>(row(0), w0, words, r);
  for (int d = 0; d < dn; d++) {
    load_words<V, kMode>(row(min(d + 1, dn - 1)), w0, words, rn);
    // keeps the load ahead of the lookups; the scheduler would sink it to
    // the copy below and wait for it at once
    __builtin_amdgcn_sched_barrier(0);
    gf_lds_lookup<V, G>(acc, lds + d * G * 64, r);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < (V + 1) / 2; k++) r[k] = rn[k];
  }
}

template <int V, int G, int kMode>
__device__ __forceinline__ void gf_lds_store(const u32x4 (&acc)[G][V], uint16_t *const *__restrict__ dsts,
                                             int ndst, int g0, long w0, long words, int accumulate) {
  const auto dp = gptr(dsts);
#pragma unroll
  for (int g = 0; g < G; g++) {
#pragma unroll
    for (int j = 0; j < kTabRows; j++) {
      const int r = kTabRows * (g0 + g) + j;
      if (r < ndst) {
        uint32_t y[V];
#pragma unroll
        for (int k = 0; k < V; k++) y[k] = (acc[g][k][j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
        store_words<V, kMode>(gptr(dp[r]), w0, words, y, accumulate != 0);
      }
    }
  }
}

// One workgroup's share of a contraction: words [tile*kTpb*V, +kTpb*V) of
// output rows [8*G*rowblk, +8*G) (variant T, LDS nibble tables).  Must be
// called by the whole workgroup (it synchronises); `lds` holds D*G KiB.
// align: the byte alignment (2, 4, 8 or 16) shared by every source and
// destination row; a thread whose V words lie inside the rows uses vector
// accesses when align >= 2V, dword pairs when align >= 4, and checked word
// accesses otherwise.
template <int V, int G, int D, bool kPrefetch = false>
__device__ __forceinline__ void gf_lds_tile(u32x4 *__restrict__ lds, const uint16_t *const *__restrict__ srcs,
                                            int nsrc, uint16_t *const *__restrict__ dsts, int ndst,
                                            const uint64_t *__restrict__ tabs, long words, int accumulate,
                                            int tile, int rowblk, int align) {
  static_assert(!kPrefetch || D <= 64, "one row pointer per lane of a wave");
  const int tid = threadIdx.x;
  const int ngroups = gf_tab_groups(ndst);
  const int g0 = rowblk * G;
  const long w0 = ((long)tile * kTpb + tid) * V;
  const bool inside = w0 + V <= words;
  const int mode = (inside && align >= 2 * V) ? 0 : (inside && align >= 4 && V % 2 == 0) ? 1 : 2;
  const P2H_GLOBAL u32x4 *gtab = reinterpret_cast<const P2H_GLOBAL u32x4 *>(gptr(tabs));
  constexpr int kMode1 = (V % 2 == 0) ? 1 : 2;

  u32x4 acc[G][V];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int k = 0; k < V; k++) acc[g][k] = (u32x4){0u, 0u, 0u, 0u};

  for (int d0 = 0; d0 < nsrc; d0 += D) {
    const int dn = min(D, nsrc - d0);
    __syncthreads();
    // stage the tables of sources d0..d0+dn-1 for this workgroup's groups:
    // lds[(d * G + g) * 64 + e] = entry e of (group g0+g, source d0+d)
    const int units = G * dn * 64;
    for (int u = tid; u < units; u += kTpb) {
      const int g = u / (dn * 64), rem = u - g * dn * 64, d = rem >> 6, e = rem & 63;
      u32x4 val = {0u, 0u, 0u, 0u};
      if (g0 + g < ngroups) val = gtab[((long)(g0 + g) * nsrc + d0 + d) * 64 + e];
      lds[(d * G + g) * 64 + e] = val;
    }
    __syncthreads();
    if constexpr (kPrefetch) {
      // row pointers of the stage, one per lane, while all lanes are active
      const int lane = tid & 63;
      const uint64_t pv = lane < dn ? (uint64_t)gptr(srcs)[d0 + lane] : 0;
      const uint32_t plo = (uint32_t)pv, phi = (uint32_t)(pv >> 32);
      if (w0 < words) {
        if (mode == 0)
          gf_lds_sources_pf<V, G, 0>(acc, lds, plo, phi, dn, w0, words);
        else if (mode == 1)
          gf_lds_sources_pf<V, G, kMode1>(acc, lds, plo, phi, dn, w0, words);
        else
          gf_lds_sources_pf<V, G, 2>(acc, lds, plo, phi, dn, w0, words);
      }
    } else if (w0 < words) {
      if (mode == 0)
        gf_lds_sources<V, G, 0>(acc, lds, srcs, d0, dn, w0, words);
      else if (mode == 1)
        gf_lds_sources<V, G, kMode1>(acc, lds, srcs, d0, dn, w0, words);
      else
        gf_lds_sources<V, G, 2>(acc, lds, srcs, d0, dn, w0, words);
    }
  }

  if (w0 >= words) return;
  if (mode == 0)
    gf_lds_store<V, G, 0>(acc, dsts, ndst, g0, w0, words, accumulate);
  else if (mode == 1)
    gf_lds_store<V, G, kMode1>(acc, dsts, ndst, g0, w0, words, accumulate);
  else
    gf_lds_store<V, G, 2>(acc, dsts, ndst, g0, w0, words, accumulate);
}

}  // namespace par2hip
