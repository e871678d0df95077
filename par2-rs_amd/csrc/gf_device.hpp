// gf_device.hpp -- device code shared by the single-set and batched
// contraction kernels (variant T, LDS nibble tables).
#pragma once
#include "common.hpp"

namespace par2hip {

constexpr int kTpb = 256;  // threads per workgroup
constexpr int kGpw = 4;    // output groups (of 4 rows) per workgroup -> 16 rows
constexpr int kDch = 16;   // sources per LDS table chunk

template <int V>
struct WordVec;
template <>
struct WordVec<8> { using T = uint4; };
template <>
struct WordVec<4> { using T = uint2; };
template <>
struct WordVec<2> { using T = uint32_t; };
template <>
struct WordVec<1> { using T = uint16_t; };

template <int V, bool kDynAlign = false>
__device__ inline void load_words(const uint16_t *__restrict__ p, long w0, long words, uint32_t (&x)[V]) {
  if (w0 + V <= words && (!kDynAlign || ((uintptr_t)(p + w0) & (2 * V - 1)) == 0)) {
    typename WordVec<V>::T t = *reinterpret_cast<const typename WordVec<V>::T *>(p + w0);
    const uint16_t *h = reinterpret_cast<const uint16_t *>(&t);
#pragma unroll
    for (int k = 0; k < V; k++) x[k] = h[k];
  } else if (kDynAlign && w0 + V <= words && ((uintptr_t)(p + w0) & 3) == 0) {
#pragma unroll
    for (int k = 0; k < V; k += 2) {
      const uint32_t t = *reinterpret_cast<const uint32_t *>(p + w0 + k);
      x[k] = t & 0xFFFFu;
      x[k + 1] = t >> 16;
    }
  } else {
#pragma unroll
    for (int k = 0; k < V; k++) x[k] = (w0 + k < words) ? p[w0 + k] : 0u;
  }
}

template <int V, bool kDynAlign = false>
__device__ inline void store_words(uint16_t *__restrict__ p, long w0, long words, const uint32_t (&x)[V],
                                   bool accumulate) {
  if (w0 + V <= words && (!kDynAlign || ((uintptr_t)(p + w0) & (2 * V - 1)) == 0)) {
    typename WordVec<V>::T t;
    uint16_t *h = reinterpret_cast<uint16_t *>(&t);
    if (accumulate) {
      t = *reinterpret_cast<const typename WordVec<V>::T *>(p + w0);
#pragma unroll
      for (int k = 0; k < V; k++) h[k] ^= (uint16_t)x[k];
    } else {
#pragma unroll
      for (int k = 0; k < V; k++) h[k] = (uint16_t)x[k];
    }
    *reinterpret_cast<typename WordVec<V>::T *>(p + w0) = t;
  } else {
#pragma unroll
    for (int k = 0; k < V; k++)
      if (w0 + k < words) p[w0 + k] = accumulate ? (uint16_t)(p[w0 + k] ^ x[k]) : (uint16_t)x[k];
  }
}

// One workgroup's share of a contraction: words [tile*kTpb*V, +kTpb*V) of
// output rows [16*rowblk, +16) (variant T, LDS nibble tables).  Must be called
// by the whole workgroup (it synchronises); `lds` holds 32 KiB.
// kDynAlign: pick 16-, 4- or 2-byte loads per source row at run time (rows of
// a set whose slice size is a multiple of 4 bytes but not of 16).
template <int V, bool kDynAlign>
__device__ __forceinline__ void gf_lds_tile(uint64_t *__restrict__ lds, const uint16_t *const *__restrict__ srcs,
                                            int nsrc, uint16_t *const *__restrict__ dsts, int ndst,
                                            const uint64_t *__restrict__ tabs, long words, int accumulate,
                                            int tile, int rowblk) {
  const int tid = threadIdx.x;
  const int ngroups = (ndst + 3) / 4;
  const int g0 = rowblk * kGpw;
  const long w0 = ((long)tile * kTpb + tid) * V;

  uint64_t acc[kGpw][V];
#pragma unroll
  for (int g = 0; g < kGpw; g++)
#pragma unroll
    for (int k = 0; k < V; k++) acc[g][k] = 0;

  for (int d0 = 0; d0 < nsrc; d0 += kDch) {
    const int dn = min(kDch, nsrc - d0);
    __syncthreads();
    // stage the tables of sources d0..d0+dn-1 for this workgroup's groups (16-byte units)
    const int units = kGpw * dn * 32;
    for (int u = tid; u < units; u += kTpb) {
      const int g = u / (dn * 32), rem = u - g * dn * 32, d = rem >> 5, k = rem & 31;
      uint4 val = make_uint4(0, 0, 0, 0);
      if (g0 + g < ngroups)
        val = reinterpret_cast<const uint4 *>(tabs + ((long)(g0 + g) * nsrc + d0 + d) * 64)[k];
      reinterpret_cast<uint4 *>(lds + (d * kGpw + g) * 64)[k] = val;
    }
    __syncthreads();
    if (w0 < words) {
      for (int d = 0; d < dn; d++) {
        uint32_t x[V];
        load_words<V, kDynAlign>(srcs[d0 + d], w0, words, x);
        const uint64_t *t = lds + d * kGpw * 64;
#pragma unroll
        for (int k = 0; k < V; k++) {
          const uint32_t a0 = x[k] & 15u, a1 = 16u + ((x[k] >> 4) & 15u);
          const uint32_t a2 = 32u + ((x[k] >> 8) & 15u), a3 = 48u + (x[k] >> 12);
#pragma unroll
          for (int g = 0; g < kGpw; g++) {
            const uint64_t *tg = t + g * 64;
            acc[g][k] ^= tg[a0] ^ tg[a1] ^ tg[a2] ^ tg[a3];
          }
        }
      }
    }
  }

  if (w0 >= words) return;
#pragma unroll
  for (int g = 0; g < kGpw; g++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int r = 4 * (g0 + g) + j;
      if (r < ndst) {
        uint32_t y[V];
#pragma unroll
        for (int k = 0; k < V; k++) y[k] = (uint32_t)(acc[g][k] >> (16 * j)) & 0xFFFFu;
        store_words<V, kDynAlign>(dsts[r], w0, words, y, accumulate != 0);
      }
    }
  }
}

}  // namespace par2hip
