// gf_kernels.hip -- GF(2^16) coefficient generation and the region
// multiply-accumulate ("contraction") kernels for gfx950.
//
// The contraction computes, for every output row i and word w,
//     dst_i[w] (= or ^=) XOR_j coeff[i][j] * src_j[w]
// which is the encoder's recovery accumulation (src/par2_rs.rs:816-870,
// muladd per (recovery, input) pair) and the repair's region muladd
// (src/galois/region/mod.rs:90-131 driven by src/par2_rs.rs:527-603) done as
// one pass over the inputs with the outputs held in registers.
//
// Variant T ("LDS nibble tables"): every coefficient c is expanded into four
// 16-entry tables T_i[v] = c * (v << 4i); four outputs are packed into one
// 64-bit entry, so one ds_read_b64 yields four partial products.  A 16-entry
// table of 8-byte entries covers all 32 LDS banks exactly once, so any lane
// pattern is conflict-free (equal addresses broadcast).
#include "common.hpp"

namespace par2hip {

// --------------------------------------------------------------- helpers --
__global__ void k_coeff_from_logs(uint16_t *__restrict__ coeff, const uint32_t *__restrict__ src_log,
                                  const uint32_t *__restrict__ exps, int nrow, int nsrc,
                                  const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nrow * nsrc) return;
  const int r = (int)(idx / nsrc), d = (int)(idx % nsrc);
  // c_d^e = 2^(n_d * e) with e already reduced mod 65535 (as gf_pow does,
  // core.rs:256): n_d * e < 2^32, folded twice into [0, 65535]
  uint32_t l = src_log[d] * exps[r];
  l = (l & 0xFFFFu) + (l >> 16);
  l = (l & 0xFFFFu) + (l >> 16);
  coeff[idx] = exp_tab[l];
}

int32_t launch_coeff_from_logs(par2hip_ctx *ctx, hipStream_t s, uint16_t *coeff,
                               const uint32_t *d_src_log, const uint32_t *d_exps, int nrow,
                               int nsrc) {
  const long total = (long)nrow * nsrc;
  if (total == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "coeff_from_logs", s);
  hipLaunchKernelGGL(k_coeff_from_logs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     coeff, d_src_log, d_exps, nrow, nsrc, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// ------------------------------------------------- variant T: nibble tables --
// Table layout: tabs[g][d][i*16 + v], g = output group (rows 4g..4g+3),
// d = source, i = nibble position, v = nibble value; entry packs the four
// rows' products c_{4g+j,d} * (v << 4i) at bits [16j, 16j+16).
__global__ void k_nibble_tables(uint64_t *__restrict__ tabs, const uint16_t *__restrict__ coeff,
                                int nrow, int nsrc, int ldc, const uint16_t *__restrict__ log_tab,
                                const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ngroups = (nrow + 3) / 4;
  if (idx >= (long)ngroups * nsrc * 64) return;
  const int iv = (int)(idx & 63);
  const long gd = idx >> 6;
  const int d = (int)(gd % nsrc), g = (int)(gd / nsrc);
  const int i = iv >> 4, v = iv & 15;
  const uint32_t x = (uint32_t)v << (4 * i);
  uint64_t e = 0;
  if (x) {
    const uint32_t lx = log_tab[x];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int r = 4 * g + j;
      if (r < nrow) {
        const uint16_t c = coeff[(long)r * ldc + d];
        if (c) e |= (uint64_t)exp_tab[lx + log_tab[c]] << (16 * j);
      }
    }
  }
  tabs[idx] = e;
}

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

template <int V>
__device__ inline void load_words(const uint16_t *__restrict__ p, long w0, long words, uint32_t (&x)[V]) {
  if (w0 + V <= words) {
    typename WordVec<V>::T t = *reinterpret_cast<const typename WordVec<V>::T *>(p + w0);
    const uint16_t *h = reinterpret_cast<const uint16_t *>(&t);
#pragma unroll
    for (int k = 0; k < V; k++) x[k] = h[k];
  } else {
#pragma unroll
    for (int k = 0; k < V; k++) x[k] = (w0 + k < words) ? p[w0 + k] : 0u;
  }
}

template <int V>
__device__ inline void store_words(uint16_t *__restrict__ p, long w0, long words, const uint32_t (&x)[V],
                                   bool accumulate) {
  if (w0 + V <= words) {
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

// Workgroup w covers words [tile*kTpb*V, +kTpb*V) of output rows
// [16*rowblk, +16).  The 1-D grid is remapped so that the row blocks of one
// word tile run on the same XCD back to back (workgroups are dealt to the 8
// XCDs round-robin), letting them share the tile's input lines in that L2.
template <int V>
__global__ __launch_bounds__(kTpb, 2) void k_gf_lds(const uint16_t *const *__restrict__ srcs, int nsrc,
                                                   uint16_t *const *__restrict__ dsts, int ndst,
                                                   const uint64_t *__restrict__ tabs, long words,
                                                   int accumulate, int ntiles, int nrowblk) {
  __shared__ uint64_t lds[kDch * kGpw * 64];  // [d][g][i*16+v], 32 KiB

  // --- XCD-aware remap of the linear workgroup id
  const int L = blockIdx.x;
  const int xcd = L & 7, q = L >> 3;
  const int tiles_per_xcd = (ntiles + 7) >> 3;
  // within an XCD: consecutive q walk the row blocks of one tile, then the next tile
  const int local_tile = q / nrowblk, rowblk = q % nrowblk;
  const int tile = local_tile * 8 + xcd;
  if (local_tile >= tiles_per_xcd || tile >= ntiles) return;

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
        load_words<V>(srcs[d0 + d], w0, words, x);
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
        store_words<V>(dsts[r], w0, words, y, accumulate != 0);
      }
    }
  }
}

template <int V>
static int32_t launch_lds_variant(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p,
                                  const uint64_t *tabs) {
  const long per_wg = (long)kTpb * V;
  const int ntiles = (int)((p.words + per_wg - 1) / per_wg);
  const int ngroups = (p.ndst + 3) / 4;
  const int nrowblk = (ngroups + kGpw - 1) / kGpw;
  const int tiles_per_xcd = (ntiles + 7) / 8;
  const long nwg = 8L * tiles_per_xcd * nrowblk;
  if (nwg > 0x7fffffffL) return fail(ctx, PAR2HIP_EINVAL, "GF problem too large for one launch");
  ScopedTiming t(ctx, p.tag ? p.tag : "gf_muladd_lds", s);
  hipLaunchKernelGGL(k_gf_lds<V>, dim3((unsigned)nwg), dim3(kTpb), 0, s, p.srcs, p.nsrc, p.dsts,
                     p.ndst, tabs, (long)p.words, p.accumulate ? 1 : 0, ntiles, nrowblk);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

int32_t run_gf_lds(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p) {
  const int ngroups = (p.ndst + 3) / 4;
  const size_t tab_bytes = (size_t)ngroups * p.nsrc * 64 * sizeof(uint64_t);
  void *tabs = nullptr;
  int32_t rc = scratch(ctx, kScratchTables, tab_bytes, &tabs);
  if (rc) return rc;
  {
    const long total = (long)ngroups * p.nsrc * 64;
    ScopedTiming t(ctx, "nibble_tables", s);
    hipLaunchKernelGGL(k_nibble_tables, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (uint64_t *)tabs, p.coeff, p.ndst, p.nsrc, p.ldc ? p.ldc : p.nsrc, ctx->d_log, ctx->d_exp);
    P2H_CHECK(ctx, hipGetLastError());
  }
  // words per thread: the widest the alignment allows (per-workgroup table
  // staging is amortised over more words); narrower only when the grid would
  // not even give every resident workgroup slot (2 per CU) one workgroup
  const int nrowblk = (ngroups + kGpw - 1) / kGpw;
  const long slots = 2L * ctx->num_cus;
  auto nwg = [&](int v) { return (long)((p.words + (size_t)kTpb * v - 1) / ((size_t)kTpb * v)) * nrowblk; };
  int v = p.align_words % 8 == 0 ? 8 : p.align_words % 4 == 0 ? 4 : p.align_words % 2 == 0 ? 2 : 1;
  while (v > 1 && nwg(v) < slots) v /= 2;
  if (v == 8) return launch_lds_variant<8>(ctx, s, p, (const uint64_t *)tabs);
  if (v == 4) return launch_lds_variant<4>(ctx, s, p, (const uint64_t *)tabs);
  if (v == 2) return launch_lds_variant<2>(ctx, s, p, (const uint64_t *)tabs);
  return launch_lds_variant<1>(ctx, s, p, (const uint64_t *)tabs);
}

int32_t run_gf(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p) {
  if (p.ndst <= 0 || p.words == 0) return PAR2HIP_OK;
  if (p.nsrc <= 0) {
    if (p.accumulate) return PAR2HIP_OK;
    return fail(ctx, PAR2HIP_EINVAL, "contraction with no sources and overwrite semantics");
  }
  // 1 = LDS tables, 2 = bit-sliced MFMA (needs 16-byte aligned operands), 0 = automatic
  const bool mfma_ok = p.align_words >= 8;
  int which = ctx->gf_kernel;
  if (which == 0) which = kAutoGfKernel;
  if (which == 2 && mfma_ok) return run_gf_mfma(ctx, s, p);
  return run_gf_lds(ctx, s, p);
}

}  // namespace par2hip
