// gf_mfma.hip -- bit-sliced GF(2^16) contraction on the gfx950 matrix cores.
//
// Multiplication by a constant c is GF(2)-linear: bit j of c*x is the parity
// of sum_i M_c[j][i] x_i with M_c[j][i] = bit j of (c * 2^i).  The whole
// contraction dst_r = XOR_d coeff[r][d] * src_d is therefore one GF(2) matrix
// product
//     C[(r,j)][w] = parity( sum_{d,i} A[(r,j)][(d,i)] * B[(d,i)][w] )
// with A = the stacked bit matrices (16R x 16N) and B = the bits of the input
// words (16N x W).  Integer sums of 0/1 products are exact in fp32 up to
// 2^24 > 16N, so the product runs on v_mfma_scale_f32_32x32x64_f8f6f4 with
// FP4 (e2m1) operands whose products are 0 or exactly 1.0, and the parity is
// read off the fp32 accumulator at the end.
//
// One 32x32x64 MFMA covers 32 A-rows = 2 output slices x 16 bits, 64 K = 4
// sources x 16 bits, and 32 words.  Lane layout (checked on the device):
//   A/B operand lane l: row/column l%32, K half l/32 (same nibble order in A
//   and B, so the K order is ours to choose);  C lane l, accumulator q:
//   column l%32, row 8*(q/4) + 4*(l/32) + q%4.
// A-rows are ordered so that lane l's 16 accumulators of a fragment are the
// 16 bits of ONE output slice (output 2*frag + l/32): row m <-> (h, j) with
// h = (m/4)%2, j = 4*(m/8) + m%4, so accumulator q is bit q.
//
// K order inside one 64-wide step (sources 4ks..4ks+3): lane half g holds
// sources 4ks+2g+s, s in {0,1}; with y = x_s0 | x_s1 << 16, nibble p of
// operand register t carries bit 4*(p%4)+t of source s = p/4.  B leaves each
// bit where it lies in its nibble as far as e2m1 allows: registers
// y & 0x11111111 (a set bit reads 0.5), y & 0x22222222 (1.0), y & 0x44444444
// (2.0) and (y >> 1) & 0x44444444 (bit 3 is the sign, so it moves to 2.0):
// five VALU operations per operand.  A is pre-packed in the same shape (one
// u32 per lane per fragment per step) and expanded on chip to the reciprocal
// values 2.0, 1.0, 0.5, 0.5, so every product of two set bits is exactly 1.0.
//
// The int8 form runs the same contraction on v_mfma_i32_32x32x32_i8 (0/1
// bytes, int32 accumulators, parity of the integer sum): a 64-wide step is
// two MFMAs, half hh covering sources 4ks+2hh and 4ks+2hh+1.  With y the
// packed pair (bit 16s+i = bit i of source s), lane half g's operand register
// t is (y >> (4g+t)) & 0x01010101: byte b carries bit 8b+4g+t of y, so the 32
// K values of one MFMA are the 32 bits of y.  The packed A of half hh is the
// FP4 form's lane m + 32hh, so both forms share k_pack_a.  At the dense rates
// (int8 5 POP/s, FP4 10 PFLOP/s) the int8 form needs twice the matrix-core
// time of the FP4 one; both are kept and measured.
#include <type_traits>

#include "common.hpp"

namespace par2hip {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kMfWaves = 8;              // waves per workgroup (two per SIMD)
constexpr int kMfThreads = 64 * kMfWaves;
// Workgroup tile shapes: the 8 waves form kWR row groups x 8/kWR column
// groups; each wave holds kRows A fragments (output pairs) x kCols B
// fragments (32 words), kKc K-steps (4 sources each) per LDS stage.  A wave
// holds 8 accumulator tiles (128 registers, one workgroup per CU) or 4 (64
// registers, 48 KiB of LDS: two workgroups per CU, so one workgroup's barrier
// and staging overlap the other's MFMAs).
//   wide  (1, 4x2, 8): 8 outputs x 512 words, 32 sources per stage; the int8
//         form's only shape
//   small (1, 2x2, 4): 4 outputs x 512 words, two workgroups per CU: 3 or
//         4 outputs fill its rows, so no MFMA multiplies padded zero rows,
//         and each input word is read by one workgroup only.  Staged form
//         (16 sources per stage, loads two stages ahead): 1 GiB at R=4 in
//         0.287-0.295 ms against 0.305-0.306 for a one-workgroup 4 x 1024
//         tile on the same device, 300 x 64 KiB at R=3 in 0.055 against
//         0.082 ms; the ring form (32 sources per stage, k_gf_mfma_ring
//         below) runs its full tiles: 1 GiB at R=4 in 0.213-0.234 ms
//   single (1, 1x4, 4): 2 outputs x 1024 words, two workgroups per CU: one
//         or two outputs (one A fragment), so no MFMA multiplies a padded
//         fragment.  1 GiB at R=1 or 2 in 0.187 ms (72% of 8 TB/s) against
//         0.27-0.28 ms with the small tile (staged form); 0.179-0.190 ms in
//         the ring form, which runs its full tiles
//   split (2, 2x4, 8): 8 outputs x 512 words; a wave reads 2 A fragments per
//         K-step instead of 4, half the LDS traffic of A per MFMA.  Alone the
//         fastest: 1024 x 1 MiB, R=102 in 5.04-5.17 ms in the staged form
//         before the five-operation expansion (wide 5.71 ms), 2000 x 512 KiB,
//         R=100 in 4.89-4.98 ms (wide 5.46 ms); a 4-tile form at two
//         workgroups per CU took 5.23 ms.  Its full tiles run in the ring
//         form: 4.26-4.33 ms and 4.16-4.22 ms
//   tall  (4, 2x4, 4): 16 outputs x 256 words, 16 sources per stage; the
//         slowest alone (6.20 ms on 1024 x 1 MiB, R=102) but the one that
//         draws the least from the MD5 chains beside it: that encode step
//         9.49 ms (wide 9.54 ms, split 9.93 ms, the 4-tile split at two
//         workgroups per CU 10.6 ms: the chains slow with the contraction's
//         clock and issue demand)
constexpr int kMfRows = 4, kMfCols = 2, kMfKc = 8;
constexpr uint32_t kNib1 = 0x11111111u, kNib2 = 0x22222222u, kNib4 = 0x44444444u;

__device__ __forceinline__ uint16_t mul_x(uint16_t v) {
  return (uint16_t)((v << 1) ^ ((v & 0x8000u) ? (kPoly & 0xFFFFu) : 0u));
}

// Apk[mf][ks][lane]: the A operand of fragment mf (outputs 2mf, 2mf+1) at
// K-step ks, packed as two 16-bit rows of the transposed bit matrices.
// perm: bits 0 and 2 of every nibble swapped (the ring form's packing, so
// that each wave expands A with five operations, see expand_fp4_a_ring).
__global__ void k_pack_a(uint32_t *__restrict__ apk, const uint16_t *__restrict__ coeff, int nrow,
                         int nsrc, int ldc, int nmf, int nks, int perm) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nmf * nks * 64) return;
  const int lane = (int)(idx & 63);
  const long fk = idx >> 6;
  const int ks = (int)(fk % nks), mf = (int)(fk / nks);
  const int m = lane & 31, g = lane >> 5;
  const int h = (m >> 2) & 1, j = 4 * (m >> 3) + (m & 3);
  const int r = 2 * mf + h;
  uint32_t out = 0;
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const int d = 4 * ks + 2 * g + s;
    uint32_t t = 0;
    if (r < nrow && d < nsrc) {
      uint16_t v = coeff[(long)r * ldc + d];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        t |= (uint32_t)((v >> j) & 1u) << i;
        v = mul_x(v);
      }
    }
    out |= t << (16 * s);
  }
  if (perm) out = (out & 0xAAAAAAAAu) | ((out & kNib1) << 2) | ((out >> 2) & kNib1);
  apk[idx] = out;
}

// int8 operand of lane half g from a packed pair y (bit 16s+i = bit i of
// source s): register t holds bits 8b+4g+t of y as bytes b = 0..3.
__device__ __forceinline__ v4i expand_i8(uint32_t y, int g) {
  const uint32_t z = y >> (4 * g);
  v4i o;
  o[0] = (int)(z & 0x01010101u);
  o[1] = (int)((z >> 1) & 0x01010101u);
  o[2] = (int)((z >> 2) & 0x01010101u);
  o[3] = (int)((z >> 3) & 0x01010101u);
  return o;
}

// B operand (input bits) of a packed pair y: e2m1 values 0.5, 1.0, 2.0, 2.0
// for bits 4p, 4p+1, 4p+2, 4p+3 of y
__device__ __forceinline__ v8i expand_fp4(uint32_t y) {
  v8i o = {0, 0, 0, 0, 0, 0, 0, 0};
  o[0] = (int)(y & kNib1);
  o[1] = (int)(y & kNib2);
  o[2] = (int)(y & kNib4);
  o[3] = (int)((y >> 1) & kNib4);
  return o;
}

// A operand (bit-matrix entries) of a packed pair y: the reciprocals 2.0,
// 1.0, 0.5, 0.5 of expand_fp4's values, so each product is 0 or 1.0
__device__ __forceinline__ uint4 expand_fp4_a(uint32_t y) {
  return make_uint4((y << 2) & kNib4, y & kNib2, (y >> 2) & kNib1, (y >> 3) & kNib1);
}

// The same A operand from the ring form's packing (bits 0 and 2 of each
// nibble swapped by k_pack_a): five operations instead of seven.
__device__ __forceinline__ v8i expand_fp4_a_ring(uint32_t y) {
  v8i o = {0, 0, 0, 0, 0, 0, 0, 0};
  o[0] = (int)(y & kNib4);
  o[1] = (int)(y & kNib2);
  o[2] = (int)(y & kNib1);
  o[3] = (int)((y >> 3) & kNib1);
  return o;
}

// Workgroup: 8 outputs (4 A fragments) x 512 words (8 waves x 2 B fragments).
// Sources are staged 32 at a time: the 32 KiB slab of input words and the
// packed A of 8 K-steps go global -> registers -> LDS one stage ahead of the
// MFMAs that consume them (one barrier per 64 MFMAs per wave).  A is shared
// by every wave, so it is expanded to FP4 once, by the staging threads; each
// wave expands only its own B columns.  Two waves per SIMD let one wave's LDS
// reads and expansion overlap the other's MFMAs.
// Lane (n, g) of wave v owns words tile + 64v + 2n + f for B fragment f.
// kI8: int8 operands (A kept packed in LDS and expanded by each wave) instead
// of FP4 (A expanded once by the staging threads).  kRagged: the launch covers
// tiles tile0.. whose words may run past the end of the rows (only the last
// tile column), loaded word by word with checks; the main launch has no such
// code, so its loads carry no branches.  The second launch bound is waves per
// SIMD (HIP passes it as amdgpu_waves_per_eu): 2 for the 8-tile shapes (one
// 8-wave workgroup per CU), 4 for the 4-tile shape (two), which caps its
// registers at 128.
template <bool kI8, bool kRagged, int kWR = 1, int kRows = kMfRows, int kCols = kMfCols, int kKc = kMfKc,
          int kAhead = 1>
__global__ __launch_bounds__(kMfThreads, kRows * kCols == 4 ? 4 : 2) void k_gf_mfma(const uint16_t *const *__restrict__ srcs, int nsrc,
                                                           uint16_t *const *__restrict__ dsts, int ndst,
                                                           const uint32_t *__restrict__ apk, int nmf, int nks,
                                                           long words, int accumulate, int ntiles, int nrowblk,
                                                           int tile0) {
  // the 8 waves form kWR row groups x kWC column groups; each wave computes
  // kRows A fragments x kCols B fragments
  constexpr int kWC = kMfWaves / kWR;
  constexpr int kFrags = kWR * kRows;              // A fragments (output pairs) per workgroup
  constexpr int kWgWords = kWC * kCols * 32;       // words per workgroup
  constexpr int kUnits = kWgWords / 8;             // 16-byte units of one source row of the tile
  constexpr int kSrcPass = kMfThreads / kUnits;    // sources one pass of the threads stages
  constexpr int kSrcStage = 4 * kKc;               // sources per stage
  constexpr int kPasses = kSrcStage / kSrcPass;
  constexpr int kSrcWave = kUnits >= 64 ? 1 : 64 / kUnits;  // sources one wave stages per pass
  constexpr int kAWords = kKc * kFrags * 64;                // packed A words per stage
  constexpr int kApt = kAWords >= kMfThreads ? kAWords / kMfThreads : 1;  // per staging thread
  constexpr bool kAPart = kAWords < kMfThreads;             // only threads tid < kAWords stage A
  static_assert((kRows * kCols == 8 || kRows * kCols == 4) && kWR * kWC == kMfWaves && (kUnits == 32 || kUnits % 64 == 0) &&
                    kPasses * kSrcPass == kSrcStage && kPasses >= 1 && kPasses <= 4 &&
                    (kApt == 1 || kApt == 2 || kApt == 4) && kApt * kMfThreads >= kAWords,
                "tile shape");
  using AUnit = typename std::conditional<kI8, uint32_t, uint4>::type;
  using Acc = typename std::conditional<kI8, v16i, v16f>::type;
  // source rows of the 8-tile shapes padded by 32 bytes, so that the rows two
  // apart that the two lane halves of a B read address do not start on the
  // same LDS bank: measured faster (C2 alone 5.04-5.07 -> 4.99-5.01 ms, the
  // tall shape in the encode step 7.14 -> 6.93 ms), though the bank-conflict
  // counter stays at 16-17%, so not all of the gain is that; the two-workgroup
  // shapes, within 1-2% either way, keep their LDS
  constexpr int kPad = kRows * kCols == 8 ? 2 : 0;
  // one LDS block, B first: the B reads of a K-step then take their row
  // offsets as instruction immediates (below 64 KiB) instead of VALU adds
  constexpr int kBBytes = 2 * kSrcStage * (kUnits + kPad) * 16;
  constexpr int kABytes = 2 * kKc * kFrags * 64 * (int)sizeof(AUnit);
  __shared__ uint4 smem[(kBBytes + kABytes) / 16];
  auto &sB = *reinterpret_cast<uint4(*)[2][kSrcStage][kUnits + kPad]>(smem);             // [buf][source][16-byte unit]
  auto &sA = *reinterpret_cast<AUnit(*)[2][kKc][kFrags][64]>(smem + kBBytes / 16);  // [buf][ks][fragment][lane]

  const int L = blockIdx.x;
  const int xcd = L & 7, q = L >> 3;
  const int tiles_per_xcd = (ntiles + 7) >> 3;
  const int local_tile = q / nrowblk, rowblk = q % nrowblk;
  const int tile_local = local_tile * 8 + xcd;
  if (local_tile >= tiles_per_xcd || tile_local >= ntiles) return;
  const int tile = tile0 + tile_local;

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int n = lane & 31, g = lane >> 5;
  const int wr = wave / kWC, wc = wave % kWC;
  const int mf0 = rowblk * kFrags;
  const long tile_w0 = (long)tile * kWgWords;
  const int nchunks = (nks + kKc - 1) / kKc;

  // staging roles.  B: unit tid + 512k -> source tid/kUnits + kSrcPass*k,
  // 16-byte column tid%kUnits.  A: thread tid owns packed words
  // kApt*tid..+kApt-1 of the stage, ordered (fragment, K-step, lane).
  //
  // The loads of stage c+kAhead are issued before stage c's MFMAs and consumed
  // only by a later store_stage, so they must not be waited for in between: no
  // load sits behind a branch or feeds a select before store_stage.  The
  // source pointers of a stage are wave-uniform (one or two per wave and
  // pass, the lane half picking between two) and fetched one stage ahead;
  // sources past the last one and A words past the last fragment or K-step
  // are loaded from clamped (valid) addresses and zeroed in store_stage.  Only
  // a ragged last word tile takes the checked per-word path.
  const int wsrc = __builtin_amdgcn_readfirstlane(wave * 64 / kUnits);  // the wave's first staging source
  const int lsrc = kSrcWave == 2 ? (lane >> 5) : 0;                    // this lane's source within the wave
  const P2H_GLOBAL uint16_t *pn[kPasses][kSrcWave];  // source rows of the next stage to load
  auto fetch_ptrs = [&](int c) {
#pragma unroll
    for (int k = 0; k < kPasses; k++)
#pragma unroll
      for (int j = 0; j < kSrcWave; j++)
        pn[k][j] = gptr(gptr(srcs)[min(kSrcStage * c + wsrc + j + kSrcPass * k, nsrc - 1)]);
  };
  using AVec = typename std::conditional<kApt == 4, u32x4, typename std::conditional<kApt == 2, u32x2, uint32_t>::type>::type;
  // kAhead stages of loads in flight: stage c's registers are set c % kAhead
  u32x4 rb[kAhead][kPasses];
  AVec ra[kAhead];
  uint32_t ok_mask[kAhead];  // bit k: this thread's source row of pass k exists; bit 8: its A words exist
  auto load_b = [&](int c, int k, uint32_t &ok) -> u32x4 {
    const P2H_GLOBAL uint16_t *src = pn[k][0];
    if constexpr (kSrcWave == 2) src = lsrc ? pn[k][1] : pn[k][0];
    const long w = tile_w0 + 8 * (tid % kUnits);
    if (kSrcStage * c + wsrc + lsrc + kSrcPass * k < nsrc) ok |= 1u << k;
    if constexpr (!kRagged) {
      return *reinterpret_cast<const P2H_GLOBAL u32x4 *>(src + w);
    } else {
      uint16_t t[8];
#pragma unroll
      for (int e = 0; e < 8; e++) t[e] = (w + e < words) ? src[w + e] : 0;
      const u32x4 v = {t[0] | (uint32_t)t[1] << 16, t[2] | (uint32_t)t[3] << 16, t[4] | (uint32_t)t[5] << 16,
                       t[6] | (uint32_t)t[7] << 16};
      return v;
    }
  };
  // the register set is a compile-time index (S), so the sets stay in registers
  auto load_stage = [&](int c, auto S) {
    uint32_t &ok = ok_mask[S];
    ok = 0;
#pragma unroll
    for (int k = 0; k < kPasses; k++) rb[S][k] = load_b(c, k, ok);
    {
      // a thread past the stage's A words (kAPart) loads word 0 again and
      // stores nothing
      const int e = (kAPart && tid >= kAWords) ? 0 : kApt * tid;
      const int frag = e / (kKc * 64), ksl = (e / 64) % kKc, l0 = e % 64;
      const int ks = kKc * c + ksl;
      if (mf0 + frag < nmf && ks < nks) ok |= 1u << 8;
      const long idx = ((long)min(mf0 + frag, nmf - 1) * nks + min(ks, nks - 1)) * 64 + l0;
      ra[S] = *reinterpret_cast<const P2H_GLOBAL AVec *>(gptr(apk) + idx);
    }
    fetch_ptrs(c + 1);
  };
  auto store_stage = [&](int buf, auto S) {
    const uint32_t ok = ok_mask[S];
    const u32x4 zero = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < kPasses; k++) {
      const u32x4 v = (ok & (1u << k)) ? rb[S][k] : zero;
      sB[buf][tid / kUnits + kSrcPass * k][tid % kUnits] = make_uint4(v.x, v.y, v.z, v.w);
    }
    if (kAPart && tid >= kAWords) return;
    const int e0 = kApt * tid, frag = e0 / (kKc * 64), ksl = (e0 / 64) % kKc, l0 = e0 % 64;
    uint32_t y[kApt];
    __builtin_memcpy(y, &ra[S], sizeof y);
#pragma unroll
    for (int e = 0; e < kApt; e++) {
      if (!(ok & 256u)) y[e] = 0;
      if constexpr (kI8)
        sA[buf][ksl][frag][l0 + e] = y[e];
      else
        sA[buf][ksl][frag][l0 + e] = expand_fp4_a(y[e]);
    }
  };

  Acc acc[kRows][kCols];
#pragma unroll
  for (int i = 0; i < kRows; i++)
#pragma unroll
    for (int f = 0; f < kCols; f++)
#pragma unroll
      for (int k = 0; k < 16; k++) acc[i][f][k] = 0;

  using Set0 = std::integral_constant<int, 0>;
  using Set1 = std::integral_constant<int, (kAhead > 1 ? 1 : 0)>;
  fetch_ptrs(0);
  load_stage(0, Set0{});
  store_stage(0, Set0{});
  if (kAhead > 1 && 1 < nchunks) load_stage(1, Set1{});
  __syncthreads();
  const int bcol = wc * (kCols * 32) + kCols * n;  // this lane's first word inside the tile
  // One stage: the loads of stage c + kAhead go out, stage c's MFMAs run from
  // LDS buffer c % 2, then stage c+1 (loaded kAhead - 1 stages earlier) is
  // stored to the other buffer, which was last read in stage c-1, which every
  // wave finished before the previous barrier.  Stage k's registers are set
  // k % kAhead, a compile-time index, so the loop runs two stages per turn.
  auto stage = [&](int c, auto S) {
    constexpr int kS = decltype(S)::value;
    const int buf = c & 1;
    if (c + kAhead < nchunks) load_stage(c + kAhead, std::integral_constant<int, (kAhead > 1 ? kS : 0)>{});
    // all K-steps of the stage (K-steps past the last source carry zero
    // operands); each wave expands its own B columns, A comes pre-expanded.
    // Kept rolled: unrolled, the operands of 8 steps spill at 2 waves/SIMD.
#pragma unroll 1
    for (int ksl = 0; ksl < kKc; ksl++) {
      if constexpr (kI8) {
        static_assert(!kI8 || (kWR == 1 && kRows == kMfRows && kCols == kMfCols && kKc == kMfKc), "int8: default shape");
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
          const uint32_t x0 = reinterpret_cast<const uint32_t *>(&sB[buf][4 * ksl + 2 * hh][0])[bcol >> 1];
          const uint32_t x1 = reinterpret_cast<const uint32_t *>(&sB[buf][4 * ksl + 2 * hh + 1][0])[bcol >> 1];
          v4i b[kMfCols];
          b[0] = expand_i8(__builtin_amdgcn_perm(x1, x0, 0x05040100u), g);  // word 0: src 2hh low, 2hh+1 high
          b[1] = expand_i8(__builtin_amdgcn_perm(x1, x0, 0x07060302u), g);  // word 1
#pragma unroll
          for (int i = 0; i < kMfRows; i++) {
            const v4i a = expand_i8(sA[buf][ksl][i][n + 32 * hh], g);
#pragma unroll
            for (int f = 0; f < kMfCols; f++)
              acc[i][f] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b[f], acc[i][f], 0, 0, 0);
          }
        }
      } else {
        // words bcol..bcol+kCols-1 of sources 4ksl+2g (x0) and 4ksl+2g+1 (x1)
        uint32_t x0[kCols / 2], x1[kCols / 2];
#pragma unroll
        for (int h = 0; h < kCols / 2; h++) {
          x0[h] = reinterpret_cast<const uint32_t *>(&sB[buf][4 * ksl + 2 * g][0])[(bcol >> 1) + h];
          x1[h] = reinterpret_cast<const uint32_t *>(&sB[buf][4 * ksl + 2 * g + 1][0])[(bcol >> 1) + h];
        }
        v8i b[kCols];
#pragma unroll
        for (int h = 0; h < kCols / 2; h++) {
          b[2 * h] = expand_fp4(__builtin_amdgcn_perm(x1[h], x0[h], 0x05040100u));      // even word: src 2g low, 2g+1 high
          b[2 * h + 1] = expand_fp4(__builtin_amdgcn_perm(x1[h], x0[h], 0x07060302u));  // odd word
        }
#pragma unroll
        for (int i = 0; i < kRows; i++) {
          const uint4 t = sA[buf][ksl][wr * kRows + i][lane];
          const v8i a = v8i{(int)t.x, (int)t.y, (int)t.z, (int)t.w, 0, 0, 0, 0};
#pragma unroll
          for (int f = 0; f < kCols; f++)
            acc[i][f] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b[f], acc[i][f], 4, 4, 0, 127, 0, 127);
        }
      }
    }
    if (c + 1 < nchunks) store_stage(buf ^ 1, std::integral_constant<int, (kAhead > 1 ? 1 - kS : 0)>{});
    __syncthreads();
  };
  for (int c = 0; c < nchunks; c += 2) {
    stage(c, Set0{});
    if (c + 1 < nchunks) stage(c + 1, Set1{});
  }

  // parity of each accumulator = bit q of output slice 2*(mf0+i) + g at words w..w+kCols-1
  const long w = tile_w0 + bcol;
  if (w >= words) return;
#pragma unroll
  for (int i = 0; i < kRows; i++) {
    const int r = 2 * (mf0 + wr * kRows + i) + g;
    if (r >= ndst) continue;
    uint32_t v[kCols];
#pragma unroll
    for (int f = 0; f < kCols; f++) {
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 16; k++) word |= ((uint32_t)acc[i][f][k] & 1u) << k;  // exact integer sums
      v[f] = word;
    }
    P2H_GLOBAL uint16_t *p = gptr(gptr(dsts)[r]);
    if (w + kCols <= words) {  // rows of the matrix-core form are 16-byte aligned
#pragma unroll
      for (int h = 0; h < kCols / 2; h++) {
        uint32_t out = v[2 * h] | (v[2 * h + 1] << 16);
        if (accumulate) out ^= *reinterpret_cast<const P2H_GLOBAL uint32_t *>(p + w + 2 * h);
        *reinterpret_cast<P2H_GLOBAL uint32_t *>(p + w + 2 * h) = out;
      }
    } else {
#pragma unroll
      for (int f = 0; f < kCols; f++)
        if (w + f < words) p[w + f] = accumulate ? (uint16_t)(p[w + f] ^ v[f]) : (uint16_t)v[f];
    }
  }
}


// Ring form of the FP4 contraction: the input words and the packed A of each
// stage go global -> LDS directly (global_load_lds_dwordx4: no VGPR staging
// and no store pass) into one of two LDS stage buffers, one stage ahead of
// the MFMAs; every wave expands its own A and B operands (five VALU
// operations each).  The loads of stage c+1 go out at the start of stage c
// into the buffer stage c-1 used, which every wave finished reading before
// the barrier that ended stage c-1; at the end of stage c every wave waits
// for all of its loads (vmcnt(0): no assumption about the order in which
// the LDS loads and the row-pointer load retire) and a barrier publishes
// stage c+1.  The two buffers of each role are separate LDS objects selected at
// compile time (the stage loop turns two stages at a time), so the compiler
// sees that the loads in flight do not write the buffer being read.  Full
// word tiles only: the ragged last tile column takes k_gf_mfma<.., true, ..>
// of the same tile width.
// A is packed with bits 0 and 2 of each nibble swapped (k_pack_a perm=1) and
// padded with zeros to whole stages and fragments, so no load needs a
// bounds check; source rows past the last one re-read the last row, which
// the zero A entries of those sources cancel (no e2m1 code is NaN or Inf).
template <int kWR, int kRows, int kCols, int kKc>
__global__ __launch_bounds__(kMfThreads, kRows * kCols == 4 ? 4 : 2) void k_gf_mfma_ring(
    const uint16_t *const *__restrict__ srcs, int nsrc, uint16_t *const *__restrict__ dsts, int ndst,
    const uint32_t *__restrict__ apk, int nks_pad, long words, int accumulate, int ntiles, int nrowblk) {
  constexpr int kWC = kMfWaves / kWR;
  constexpr int kFrags = kWR * kRows;         // A fragments (output pairs) per workgroup
  constexpr int kWgWords = kWC * kCols * 32;  // words per workgroup
  constexpr int kRowBytes = kWgWords * 2;     // bytes of one source row of the tile
  constexpr int kSrcStage = 4 * kKc;          // sources per stage
  constexpr int kBBytes = kSrcStage * kRowBytes;
  constexpr int kFragBytes = kKc * 256;  // packed A of one fragment for one stage
  constexpr int kABytes = kFrags * kFragBytes;
  // Loads are 64 lanes x 16 bytes (the LDS copy of a narrower load puts one
  // lane per dword, so only whole dwords per lane keep the layout packed).
  // B load j covers bytes 1024j.. of the stage's row-major B block, A load j
  // bytes 1024j.. of its fragment-major A.  Every wave issues the same number
  // of loads, with no branch (a branch around a load makes the compiler's
  // wait insertion fall back to waiting for every load in flight where the
  // paths join): when the stage's A is under 8 KiB, wave w loads A piece
  // w mod kALoads, so several waves write the same bytes to the same place.
  constexpr int kBLoads = kBBytes / 1024, kBPerWave = kBLoads / kMfWaves;
  constexpr int kALoads = kABytes / 1024, kANum = (kALoads + kMfWaves - 1) / kMfWaves;
  static_assert((kRows * kCols == 8 || kRows * kCols == 4) && kWR * kWC == kMfWaves && kRowBytes % 1024 == 0 &&
                    kBLoads % kMfWaves == 0 && kBPerWave <= 64 && kABytes % 1024 == 0 && kFragBytes % 1024 == 0 &&
                    (kALoads < kMfWaves || kALoads % kMfWaves == 0),
                "ring tile shape");
  __shared__ uint4 sB0[kBBytes / 16], sB1[kBBytes / 16];
  __shared__ uint32_t sA0[kABytes / 4], sA1[kABytes / 4];

  const int L = blockIdx.x;
  const int xcd = L & 7, q = L >> 3;
  const int tiles_per_xcd = (ntiles + 7) >> 3;
  const int local_tile = q / nrowblk, rowblk = q % nrowblk;
  const int tile = local_tile * 8 + xcd;
  if (local_tile >= tiles_per_xcd || tile >= ntiles) return;

  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int n = lane & 31, g = lane >> 5;
  const int wr = wave / kWC, wc = wave % kWC;
  const int mf0 = rowblk * kFrags;
  const long tile_w0 = (long)tile * kWgWords;
  const int nchunks = nks_pad / kKc;

  // B load k of the wave (load j = wave + 8k) reads one source row (a row
  // is one or more whole loads), at byte 1024j mod kRowBytes + 16 * lane of
  // the tile.  The rows' pointers for the next stage to issue are fetched one
  // stage before with one vector load (lane k holds load k's pointer) and
  // read back with readlane: no scalar load is in flight to turn every wait
  // on an LDS read into a wait for it, and a single register carries them
  // from stage to stage.
  auto brow = [&](int k) { return (1024 * (wave + kMfWaves * k)) / kRowBytes; };
  auto boff = [&](int k) { return (1024 * (wave + kMfWaves * k)) % kRowBytes + 16 * lane; };
  uint64_t pv = 0;
  auto fetch_ptrs = [&](int c) {
    int row = 0;
#pragma unroll
    for (int k = 0; k < kBPerWave; k++)
      if (lane == k) row = brow(k);
    const int src = min(kSrcStage * c + min(row, kSrcStage - 1), nsrc - 1);
    pv = reinterpret_cast<const P2H_GLOBAL uint64_t *>(gptr(srcs))[src];
  };
  auto issue = [&](int c, uint4 *bB, uint32_t *bA) {
    const uint16_t *p[kBPerWave];
#pragma unroll
    for (int k = 0; k < kBPerWave; k++)
      p[k] = reinterpret_cast<const uint16_t *>(
          (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pv, k) |  // zero-extended halves
          (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pv >> 32), k) << 32);
    fetch_ptrs(c + 1);
#pragma unroll
    for (int k = 0; k < kBPerWave; k++) {
      const int j = wave + kMfWaves * k;
      __builtin_amdgcn_global_load_lds((const void *)(p[k] + tile_w0 + boff(k) / 2),
                                       (__attribute__((address_space(3))) void *)(bB + 64 * j), 16, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < kANum; k++) {
      const int j = kALoads < kMfWaves ? wave % kALoads : wave + kMfWaves * k;
      const int frag = (1024 * j) / kFragBytes, off = (1024 * j) % kFragBytes;
      const uint8_t *a = reinterpret_cast<const uint8_t *>(apk) +
                         ((long)(mf0 + frag) * nks_pad + (long)kKc * c) * 256 + off + 16 * lane;
      __builtin_amdgcn_global_load_lds((const void *)a,
                                       (__attribute__((address_space(3))) void *)(bA + 256 * j), 16, 0, 0);
    }
  };

  v16f acc[kRows][kCols];
#pragma unroll
  for (int i = 0; i < kRows; i++)
#pragma unroll
    for (int f = 0; f < kCols; f++)
#pragma unroll
      for (int k = 0; k < 16; k++) acc[i][f][k] = 0;

  // wait for this wave's loads, then for every wave's (the barrier's fences
  // keep the compiler from moving LDS reads across it)
  auto publish = [&] {
    // the stage's MFMAs stay above the wait: they touch no memory, so the
    // compiler would otherwise sink them below the barrier and expose the
    // full latency of the loads just sent; an empty asm that takes one
    // element of every accumulator pins them here
#pragma unroll
    for (int i = 0; i < kRows; i++)
#pragma unroll
      for (int f = 0; f < kCols; f++) asm volatile("" ::"v"(acc[i][f][0]));
    __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8));  // vmcnt(0)
    __syncthreads();
  };
  const int bcol = wc * (kCols * 32) + kCols * n;  // this lane's first word inside the tile
  auto kloop = [&](const uint4 *bB, const uint32_t *bA) {
#pragma unroll
    for (int ksl = 0; ksl < kKc; ksl++) {
      const uint32_t *r0 = reinterpret_cast<const uint32_t *>(bB) + (4 * ksl + 2 * g) * (kRowBytes / 4);
      const uint32_t *r1 = r0 + kRowBytes / 4;
      uint32_t x0[kCols / 2], x1[kCols / 2];
#pragma unroll
      for (int h = 0; h < kCols / 2; h++) {
        x0[h] = r0[(bcol >> 1) + h];
        x1[h] = r1[(bcol >> 1) + h];
      }
      uint32_t ya[kRows];
#pragma unroll
      for (int i = 0; i < kRows; i++) ya[i] = bA[((wr * kRows + i) * kKc + ksl) * 64 + lane];
      v8i b[kCols];
#pragma unroll
      for (int h = 0; h < kCols / 2; h++) {
        b[2 * h] = expand_fp4(__builtin_amdgcn_perm(x1[h], x0[h], 0x05040100u));      // even word
        b[2 * h + 1] = expand_fp4(__builtin_amdgcn_perm(x1[h], x0[h], 0x07060302u));  // odd word
      }
#pragma unroll
      for (int i = 0; i < kRows; i++) {
        const v8i a = expand_fp4_a_ring(ya[i]);
#pragma unroll
        for (int f = 0; f < kCols; f++)
          acc[i][f] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b[f], acc[i][f], 4, 4, 0, 127, 0, 127);
      }
    }
  };
  // stage c computes from (bB, bA) after sending the loads of stage c+1 into
  // (nB, nA)
  auto stage = [&](int c, const uint4 *bB, const uint32_t *bA, uint4 *nB, uint32_t *nA) {
    if (c + 1 < nchunks) issue(c + 1, nB, nA);
    kloop(bB, bA);
    if (c + 1 < nchunks) publish();
  };
  fetch_ptrs(0);
  issue(0, sB0, sA0);
  publish();
  for (int c = 0; c < nchunks; c += 2) {
    stage(c, sB0, sA0, sB1, sA1);
    if (c + 1 < nchunks) stage(c + 1, sB1, sA1, sB0, sA0);
  }
  // every load issued (the last one is the row-pointer load for the stage
  // after the last, clamped to the last source row) was waited for by the
  // last publish; this wait only states that no load is in flight past here
  __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8));

  // parity of each accumulator = bit q of output slice 2*(mf0+i) + g at words w..w+kCols-1
  const long w = tile_w0 + bcol;
#pragma unroll
  for (int i = 0; i < kRows; i++) {
    const int r = 2 * (mf0 + wr * kRows + i) + g;
    if (r >= ndst) continue;
    uint32_t v[kCols];
#pragma unroll
    for (int f = 0; f < kCols; f++) {
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 16; k++) word |= ((uint32_t)acc[i][f][k] & 1u) << k;  // exact integer sums
      v[f] = word;
    }
    P2H_GLOBAL uint16_t *p = gptr(gptr(dsts)[r]);
#pragma unroll
    for (int h = 0; h < kCols / 2; h++) {
      uint32_t out = v[2 * h] | (v[2 * h + 1] << 16);
      if (accumulate) out ^= *reinterpret_cast<const P2H_GLOBAL uint32_t *>(p + w + 2 * h);
      *reinterpret_cast<P2H_GLOBAL uint32_t *>(p + w + 2 * h) = out;
    }
  }
}

int32_t run_gf_mfma(par2hip_ctx *ctx, hipStream_t s, const GfProblem &p, bool int8) {
  const int nmf = (p.ndst + 1) / 2;
  const int nks = (p.nsrc + 3) / 4;
  // tile shape: 0 wide, 1 small, 2 split, 3 tall, 4 single (the table above)
  const int shape = int8 ? 0 : nmf == 1 ? 4 : nmf == 2 ? 1 : p.beside_chains ? 3 : 2;
  static const int kFragsOf[5] = {kMfRows, 2, 4, 8, 1};
  static const long kWordsOf[5] = {kMfWaves * kMfCols * 32, 512, 512, 256, 1024};
  const int frags = kFragsOf[shape];
  const long wg_words = kWordsOf[shape];
  const int ntiles = (int)((p.words + wg_words - 1) / wg_words);
  const int nfull = (int)(p.words / wg_words);  // tiles whose words all lie inside the rows
  const int nrowblk = (nmf + frags - 1) / frags;
  const long nwg = 8L * ((ntiles + 7) / 8) * nrowblk;
  if (nwg > 0x7fffffffL) return fail(ctx, PAR2HIP_EINVAL, "GF problem too large for one launch");
  // the single, small and split shapes run their full tiles in the ring
  // form, whose packed A is permuted and padded to whole fragments and
  // stages; the other shapes and the ragged tile column take the plain
  // packing.  The tall shape (beside the
  // MD5 chains) keeps the staged form: its ring form was faster (the 1 GiB,
  // R=102 contraction on 224 CUs in 5.83-5.89 ms against 6.65-6.73) but drew
  // the shader clock down under the chains, so the encode step slowed (MD5
  // 9.46-9.54 -> 10.06-10.13 ms, step 9.57-9.65 -> 10.21-10.27 ms, same box)
  const int ring_kc = shape == 4 ? 4 : 8;  // K-steps per stage: two stage buffers of each role fit the LDS
  const bool ring = !int8 && (shape == 1 || shape == 2 || shape == 4) && nfull > 0;
  const bool plain = !ring || ntiles > nfull;
  const int nmf_pad = nrowblk * frags, nks_pad = (nks + ring_kc - 1) / ring_kc * ring_kc;
  const size_t apk_bytes = plain ? (size_t)nmf * nks * 64 * sizeof(uint32_t) : 0;
  const size_t ring_bytes = ring ? (size_t)nmf_pad * nks_pad * 64 * sizeof(uint32_t) : 0;
  void *buf = nullptr;
  int32_t rc = scratch(ctx, kScratchTables, apk_bytes + ring_bytes, &buf);
  if (rc) return rc;
  uint32_t *apk = (uint32_t *)buf, *apk_ring = (uint32_t *)((char *)buf + apk_bytes);
  {
    ScopedTiming t(ctx, "pack_bitmatrices", s);
    const int ldc = p.ldc ? p.ldc : p.nsrc;
    if (plain) {
      const long total = (long)nmf * nks * 64;
      hipLaunchKernelGGL(k_pack_a, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, apk, p.coeff, p.ndst,
                         p.nsrc, ldc, nmf, nks, 0);
    }
    if (ring) {
      const long total = (long)nmf_pad * nks_pad * 64;
      hipLaunchKernelGGL(k_pack_a, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, apk_ring, p.coeff,
                         p.ndst, p.nsrc, ldc, nmf_pad, nks_pad, 1);
    }
    P2H_CHECK(ctx, hipGetLastError());
  }
  ScopedTiming t(ctx, p.tag ? p.tag : (int8 ? "gf_muladd_mfma_i8" : "gf_muladd_mfma"), s);
  auto launch = [&](auto kernel, int tiles, int tile0) {
    const long grid = 8L * ((tiles + 7) / 8) * nrowblk;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(kMfThreads), 0, s, p.srcs, p.nsrc, p.dsts, p.ndst,
                       (const uint32_t *)apk, nmf, nks, (long)p.words, p.accumulate ? 1 : 0, tiles, nrowblk, tile0);
  };
  auto launch_ring = [&](auto kernel) {
    const long grid = 8L * ((nfull + 7) / 8) * nrowblk;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(kMfThreads), 0, s, p.srcs, p.nsrc, p.dsts, p.ndst,
                       (const uint32_t *)apk_ring, nks_pad, (long)p.words, p.accumulate ? 1 : 0, nfull, nrowblk);
  };
  if (int8) {
    if (nfull) launch(k_gf_mfma<true, false>, nfull, 0);
    if (ntiles > nfull) launch(k_gf_mfma<true, true>, ntiles - nfull, nfull);
  } else if (shape == 1) {
    if (nfull) launch_ring(k_gf_mfma_ring<1, 2, 2, 8>);
    if (ntiles > nfull) launch(k_gf_mfma<false, true, 1, 2, 2, 4>, ntiles - nfull, nfull);
  } else if (shape == 4) {
    if (nfull) launch_ring(k_gf_mfma_ring<1, 1, 4, 4>);
    if (ntiles > nfull) launch(k_gf_mfma<false, true, 1, 1, 4, 4>, ntiles - nfull, nfull);
  } else if (shape == 2) {
    if (nfull) launch_ring(k_gf_mfma_ring<2, 2, 4, 8>);
    if (ntiles > nfull) launch(k_gf_mfma<false, true, 2, 2, 4, 8>, ntiles - nfull, nfull);
  } else if (shape == 3) {
    if (nfull) launch(k_gf_mfma<false, false, 4, 2, 4, 4>, nfull, 0);
    if (ntiles > nfull) launch(k_gf_mfma<false, true, 4, 2, 4, 4>, ntiles - nfull, nfull);
  } else {
    if (nfull) launch(k_gf_mfma<false, false>, nfull, 0);
    if (ntiles > nfull) launch(k_gf_mfma<false, true>, ntiles - nfull, nfull);
  }
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

}  // namespace par2hip
