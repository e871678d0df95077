// gj_kernels.hip -- device Gauss-Jordan inversion of the erasure matrix and
// the fold of the inverse into one contraction matrix.
//
// The reference eliminates on the host once per repair
// (compute_reconstruction_transform, src/par2_rs.rs:319-422) and then replays
// the recorded row operations on every chunk of every rebuilt slice
// (par2_rs.rs:618-655).  Here the inverse A^-1 is formed on the device and
// folded: missing = A^-1 * recovery  XOR  (A^-1 V) * present, i.e. a single
// m x n coefficient matrix over [present data | recovery] that the contraction
// kernel applies in one pass -- no per-chunk replay.
#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace par2hip {

// log of c^e for a constant with log n (n < 65535) and an exponent reduced
// mod 65535 on the host: n*e < 2^32, folded twice to a value in [0, 65535]
// that indexes the doubled exp table like a reduced logarithm.
__device__ __forceinline__ uint32_t fold_mod65535(uint32_t x) {
  x = (x & 0xFFFFu) + (x >> 16);
  return (x & 0xFFFFu) + (x >> 16);
}

__global__ void k_build_aug(uint16_t *__restrict__ aug, const uint32_t *__restrict__ miss_log,
                            const uint32_t *__restrict__ exps, int m,
                            const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)m * 2 * m) return;
  const int i = (int)(idx / (2 * m)), j = (int)(idx % (2 * m));
  uint16_t v;
  if (j < m) {
    // A[i][j] = c_{missing_j}^{e_i}  (par2_rs.rs:347-359)
    v = exp_tab[fold_mod65535(miss_log[j] * exps[i])];
  } else {
    v = (j - m == i) ? 1 : 0;
  }
  aug[idx] = v;
}

int32_t launch_build_aug(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, const uint32_t *d_miss_log,
                         const uint32_t *d_exps, int m) {
  const long total = (long)m * 2 * m;
  ScopedTiming t(ctx, "build_erasure_matrix", s);
  hipLaunchKernelGGL(k_build_aug, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d_aug,
                     d_miss_log, d_exps, m, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

constexpr int kGjThreads = 1024;
constexpr int kGjMaxM = 32768;      // distinct PAR2 constants
constexpr int kGjBlockedMinM = 256;  // above this the blocked multi-workgroup form runs
constexpr int kGjLdsMaxM = 180;  // [A | I] of 4*m*m bytes stays in LDS up to here
constexpr uint32_t kNoLog = 0xFFFFFFFFu;

// One workgroup.  Pivot rule as the reference: keep A[k][k] if non-zero,
// otherwise take the first row below with a non-zero entry (par2_rs.rs:362-371).
// kInLds: the m x 2m workspace is copied into LDS for the elimination.
// Elimination threads form a 32 x 32 grid over (rows, columns).
template <bool kInLds>
__global__ __launch_bounds__(kGjThreads) void k_gauss_jordan(uint16_t *__restrict__ gaug, int m,
                                                            const uint16_t *__restrict__ log_tab,
                                                            const uint16_t *__restrict__ exp_tab,
                                                            int *__restrict__ status) {
  extern __shared__ uint32_t sh[];
  const int w = 2 * m;
  uint32_t *lrow = sh;             // 2m: log of the normalised pivot row, kNoLog for zero
  uint32_t *lfac = sh + w;         // m: log of the factor of each row, kNoLog for zero / pivot row
  uint16_t *aug = kInLds ? reinterpret_cast<uint16_t *>(sh + w + m) : gaug;
  __shared__ int piv;
  const int tid = threadIdx.x;
  if (kInLds) {
    for (int e = tid; e < m * w; e += kGjThreads) aug[e] = gaug[e];
    __syncthreads();
  }
  const int tr = tid >> 5, tc = tid & 31;
  for (int k = 0; k < m; k++) {
    if (tid == 0) {
      int p = k;
      if (aug[(long)k * w + k] == 0) {
        p = m;
        for (int r = k + 1; r < m; r++)
          if (aug[(long)r * w + k] != 0) {
            p = r;
            break;
          }
      }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p == m) {
      if (tid == 0) *status = 1;
      return;
    }
    if (p != k) {
      for (int j = tid; j < w; j += kGjThreads) {
        const uint16_t t = aug[(long)k * w + j];
        aug[(long)k * w + j] = aug[(long)p * w + j];
        aug[(long)p * w + j] = t;
      }
      __syncthreads();
    }
    // normalise the pivot row by 1/pivot (exp is doubled: linv = 65535 reads 1)
    const uint32_t linv = kOrder - log_tab[aug[(long)k * w + k]];
    for (int j = tid; j < w; j += kGjThreads) {
      const uint16_t a = aug[(long)k * w + j];
      uint32_t l = kNoLog;
      if (a) {
        const uint16_t v = exp_tab[log_tab[a] + linv];
        aug[(long)k * w + j] = v;
        l = log_tab[v];
      }
      lrow[j] = l;
    }
    for (int r = tid; r < m; r += kGjThreads) {
      const uint16_t f = (r == k) ? 0 : aug[(long)r * w + k];
      lfac[r] = f ? log_tab[f] : kNoLog;
    }
    __syncthreads();
    // eliminate column k from every other row (pivot-row columns < k are zero)
    for (int r = tr; r < m; r += kGjThreads / 32) {
      const uint32_t lf = lfac[r];
      if (lf == kNoLog) continue;
      for (int j = k + tc; j < w; j += 32) {
        const uint32_t l = lrow[j];
        if (l != kNoLog) aug[(long)r * w + j] ^= exp_tab[lf + l];
      }
    }
    __syncthreads();
  }
  if (kInLds)
    for (int e = tid; e < m * w; e += kGjThreads) gaug[e] = aug[e];
  if (tid == 0) *status = 0;
}

// ---- blocked Gauss-Jordan for large m.  Panels of kGjPanel pivot columns
// K = [k, k+b): one workgroup finds the panel's pivot rows (first non-zero
// entry below, column by column, on an eliminated copy of the panel) and
// inverts their b x b block B; the row swaps are applied to all columns; then
// the pivot rows become P' = B^-1 P and every other row r takes
// row_r ^= M[r, K] * P' -- both are the region contraction (b sources, rows
// as outputs, words [k, 2m)), so the O(m^3) work runs at the contraction's
// rate.  Columns below k are already reduced (identity in the pivot rows,
// zero elsewhere) and stay untouched.  The pivots differ from the one-column
// rule only in which independent rows are taken; the inverse is unique.
constexpr int kGjPanel = 64;
constexpr int kGjCand = 2 * kGjPanel;  // rows of the fast pivot window

struct GjWork {
  int *perm;        // kGjPanel: row k+c swaps with row k+perm[c] (panel-relative)
  int *idx;         // m: panel-relative row order after the swaps
  uint16_t *q;      // (m - k) x b panel copy, column-major (column j at q + j * (m - k))
  uint16_t *binv;   // b x b
  uint16_t *newp;   // b rows of 2m words
  uint16_t *coef;   // (m - b) x b coefficients of the rank-b update
  const uint16_t **psrc;  // pivot rows + k          (b)
  uint16_t **pdst;        // newp rows               (b)
  const uint16_t **usrc;  // newp rows               (b)
  uint16_t **udst;        // the other rows + k      (m - b)
};

static size_t gj_work_bytes(int m) {
  auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t M = (size_t)m, B = kGjPanel;
  return r(B * 4) + r(M * 4) + r(M * B * 2) + r(B * B * 2) + r(B * 2 * M * 2) + r(M * B * 2) + 2 * r(B * 8) +
         r(B * 8) + r(M * 8);
}

__global__ void k_gj_init(int *status) { *status = 0; }

// One workgroup: pivots of panel [k, k+b) and the inverse of its pivot block.
__global__ __launch_bounds__(kGjThreads) void k_gj_panel(const uint16_t *__restrict__ aug, int m, int k, int b,
                                                         GjWork w, const uint16_t *__restrict__ log_tab,
                                                         const uint16_t *__restrict__ exp_tab, int *status) {
  __shared__ uint32_t lrow[kGjPanel];          // log of the normalised pivot row's entries (kNoLog: zero)
  __shared__ uint16_t blk[kGjPanel * 2 * kGjPanel];  // [B | I]
  __shared__ uint16_t fac[kGjPanel];                   // column c of [B | I] (0 in row c)
  __shared__ int piv;
  const int tid = threadIdx.x;
  const int rows = m - k;
  if (*status) {  // an earlier panel was singular: keep every access in bounds, skip the work
    if (tid < kGjPanel) w.perm[tid] = tid;
    return;
  }
  // Fast path: the pivots among the first kGjCand rows of the panel, reduced
  // in LDS.  Rows of an invertible erasure matrix are almost always
  // independent well within that window; when they are not (a structured or
  // singular matrix), the full search below runs instead.
  __shared__ uint16_t cand[kGjCand * kGjPanel];  // [row][column]
  __shared__ int cidx[kGjCand];
  __shared__ int cperm[kGjPanel];
  const int nc = rows < kGjCand ? rows : kGjCand;
  for (int e = tid; e < nc * b; e += kGjThreads) {
    const int i = e / b, j = e % b;
    cand[i * kGjPanel + j] = aug[(long)(k + i) * 2 * m + k + j];
  }
  for (int i = tid; i < nc; i += kGjThreads) cidx[i] = i;
  __syncthreads();
  bool fast = true;
  for (int c = 0; c < b; c++) {
    if (tid == 0) piv = nc;
    __syncthreads();
    if (tid >= c && tid < nc && cand[tid * kGjPanel + c] != 0) atomicMin(&piv, tid);
    __syncthreads();
    const int p = piv;
    if (p == nc) {
      fast = false;
      break;
    }
    if (p != c) {
      for (int j = c + tid; j < b; j += kGjThreads) {
        const uint16_t t = cand[c * kGjPanel + j];
        cand[c * kGjPanel + j] = cand[p * kGjPanel + j];
        cand[p * kGjPanel + j] = t;
      }
      if (tid == 0) {
        const int t = cidx[c];
        cidx[c] = cidx[p];
        cidx[p] = t;
      }
    }
    if (tid == 0) cperm[c] = p;
    __syncthreads();
    const uint32_t linv = kOrder - log_tab[cand[c * kGjPanel + c]];
    for (int j = c + 1 + tid; j < b; j += kGjThreads) {
      const uint16_t a = cand[c * kGjPanel + j];
      lrow[j] = a ? (uint32_t)log_tab[a] + linv : kNoLog;
    }
    __syncthreads();
    // rows c+1.. of the window: column c read as the factor, columns > c updated
    const int width = b - c - 1, below = nc - c - 1;
    for (int e = tid; e < width * below; e += kGjThreads) {
      const int i = c + 1 + e / width, j = c + 1 + e % width;
      const uint16_t f = cand[i * kGjPanel + c];
      const uint32_t l = lrow[j];
      if (f && l != kNoLog) cand[i * kGjPanel + j] ^= exp_tab[log_tab[f] + (l >= kOrder ? l - kOrder : l)];
    }
    __syncthreads();
  }
  if (fast) {
    for (int c = tid; c < b; c += kGjThreads) {
      w.perm[c] = cperm[c];
      w.idx[c] = cidx[c];
    }
    __syncthreads();
  } else {
  // the panel, column-major, and the identity row order
  for (long e = tid; e < (long)rows * b; e += kGjThreads) {
    const int j = (int)(e / rows), i = (int)(e % rows);
    w.q[e] = aug[(long)(k + i) * 2 * m + k + j];
  }
  for (int i = tid; i < rows; i += kGjThreads) w.idx[i] = i;
  __syncthreads();
  for (int c = 0; c < b; c++) {
    uint16_t *qc = w.q + (long)c * rows;
    if (tid == 0) piv = rows;
    __syncthreads();
    // first row at or below c with a non-zero entry in column c
    for (int i = c + tid; i < rows; i += kGjThreads)
      if (qc[i] != 0) {
        atomicMin(&piv, i);
        break;
      }
    __syncthreads();
    const int p = piv;
    if (p == rows) {
      if (tid == 0) *status = 1;
      for (int t = tid; t < kGjPanel; t += kGjThreads) w.perm[t] = t;
      return;
    }
    if (p != c) {  // swap rows c and p in the remaining columns and in the order
      for (int j = c + tid; j < b; j += kGjThreads) {
        uint16_t *qj = w.q + (long)j * rows;
        const uint16_t t = qj[c];
        qj[c] = qj[p];
        qj[p] = t;
      }
      if (tid == 0) {
        const int t = w.idx[c];
        w.idx[c] = w.idx[p];
        w.idx[p] = t;
      }
    }
    if (tid == 0) w.perm[c] = p;
    __syncthreads();
    // row c normalised by its pivot, as logs
    const uint32_t linv = kOrder - log_tab[qc[c]];
    for (int j = c + 1 + tid; j < b; j += kGjThreads) {
      const uint16_t a = w.q[(long)j * rows + c];
      lrow[j] = a ? (uint32_t)log_tab[a] + linv : kNoLog;
    }
    __syncthreads();
    // eliminate column c from the rows below (all that later pivots need)
    for (int i = c + 1 + tid; i < rows; i += kGjThreads) {
      const uint16_t f = qc[i];
      if (!f) continue;
      const uint32_t lf = log_tab[f];
      for (int j = c + 1; j < b; j++) {
        const uint32_t l = lrow[j];
        if (l != kNoLog) {
          uint32_t s = lf + (l >= kOrder ? l - kOrder : l);
          w.q[(long)j * rows + i] ^= exp_tab[s];
        }
      }
    }
    __syncthreads();
  }
  }  // full search
  // B = the pivot rows' original entries in the panel columns, then [B | I]
  // reduced in LDS
  const int bw = 2 * b;
  for (int e = tid; e < b * bw; e += kGjThreads) {
    const int i = e / bw, j = e % bw;
    blk[e] = j < b ? aug[(long)(k + w.idx[i]) * 2 * m + k + j] : (uint16_t)(j - b == i);
  }
  __syncthreads();
  for (int c = 0; c < b; c++) {
    if (tid == 0) {
      int p2 = c;
      while (p2 < b && blk[p2 * bw + c] == 0) p2++;
      piv = p2;
    }
    __syncthreads();
    const int p2 = piv;
    if (p2 == b) {  // cannot happen for independent rows; report it as singular
      if (tid == 0) *status = 1;
      return;
    }
    if (p2 != c)
      for (int j = tid; j < bw; j += kGjThreads) {
        const uint16_t t = blk[c * bw + j];
        blk[c * bw + j] = blk[p2 * bw + j];
        blk[p2 * bw + j] = t;
      }
    __syncthreads();
    // every thread reads the pivot before row c is normalised, and the other
    // rows' factors before any of them is updated
    const uint32_t linv = kOrder - log_tab[blk[c * bw + c]];
    if (tid < b) fac[tid] = tid == c ? 0 : blk[tid * bw + c];
    __syncthreads();
    if (tid < bw) {
      const uint16_t a = blk[c * bw + tid];
      blk[c * bw + tid] = a ? exp_tab[log_tab[a] + linv] : 0;
    }
    __syncthreads();
    for (int e = tid; e < b * bw; e += kGjThreads) {
      const int i = e / bw, j = e % bw;
      const uint16_t f = fac[i];
      const uint16_t v = blk[c * bw + j];
      if (f && v) blk[e] ^= exp_tab[log_tab[f] + log_tab[v]];
    }
    __syncthreads();
  }
  for (int e = tid; e < b * b; e += kGjThreads) w.binv[e] = blk[(e / b) * bw + b + e % b];
}

// The panel's row swaps, in order, on columns [k, 2m).
__global__ void k_gj_swap(uint16_t *__restrict__ aug, int m, int k, int b, const int *__restrict__ perm) {
  const int col = k + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (col >= 2 * m) return;
  for (int c = 0; c < b; c++) {
    const int p = perm[c];
    if (p != c) {
      uint16_t *x = aug + (long)(k + c) * 2 * m + col, *y = aug + (long)(k + p) * 2 * m + col;
      const uint16_t t = *x;
      *x = *y;
      *y = t;
    }
  }
}

// Coefficients and row pointers of the two contractions: P' = B^-1 P into
// newp, then row_r ^= M[r, K] P' for the m - b other rows.
__global__ void k_gj_gather(uint16_t *__restrict__ aug, int m, int k, int b, GjWork w) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int others = m - b;
  if (t < (long)others * b) {
    const int r = (int)(t / b), j = (int)(t % b);
    const int row = r < k ? r : r + b;
    w.coef[t] = aug[(long)row * 2 * m + k + j];
  }
  if (t < others) {
    const int row = (int)t < k ? (int)t : (int)t + b;
    w.udst[t] = aug + (long)row * 2 * m + k;
  }
  if (t < b) {
    w.psrc[t] = aug + (long)(k + t) * 2 * m + k;
    w.pdst[t] = w.newp + t * 2 * m;
    w.usrc[t] = w.newp + t * 2 * m;
  }
}

static int32_t gauss_jordan_blocked(par2hip_ctx *ctx, hipStream_t s, uint16_t *aug, int m, int *d_status) {
  void *ws = nullptr;
  int32_t rc;
  if ((rc = scratch(ctx, kScratchGjWork, gj_work_bytes(m), &ws))) return rc;
  auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  uint8_t *p = static_cast<uint8_t *>(ws);
  GjWork w;
  const size_t M = (size_t)m, B = kGjPanel;
  w.perm = (int *)p;
  p += r(B * 4);
  w.idx = (int *)p;
  p += r(M * 4);
  w.q = (uint16_t *)p;
  p += r(M * B * 2);
  w.binv = (uint16_t *)p;
  p += r(B * B * 2);
  w.newp = (uint16_t *)p;
  p += r(B * 2 * M * 2);
  w.coef = (uint16_t *)p;
  p += r(M * B * 2);
  w.psrc = (const uint16_t **)p;
  p += r(B * 8);
  w.pdst = (uint16_t **)p;
  p += r(B * 8);
  w.usrc = (const uint16_t **)p;
  p += r(B * 8);
  w.udst = (uint16_t **)p;
  ScopedTiming t(ctx, "gauss_jordan_blocked", s);  // the whole elimination, contractions included
  hipLaunchKernelGGL(k_gj_init, dim3(1), dim3(1), 0, s, d_status);
  P2H_CHECK(ctx, hipGetLastError());
  // every row pointer is aug + row * 2m + k with k a multiple of kGjPanel
  const int align = std::min({align_words_of((uintptr_t)aug), align_words_of((uintptr_t)w.newp),
                              align_words_of((uintptr_t)(4 * M)), align_words_of((uintptr_t)(2 * kGjPanel))});
  for (int k = 0; k < m; k += kGjPanel) {
    const int b = std::min(kGjPanel, m - k);
    const long words = 2L * m - k;
    {
      ScopedTiming tp(ctx, "gauss_jordan_pivots", s);
      hipLaunchKernelGGL(k_gj_panel, dim3(1), dim3(kGjThreads), 0, s, aug, m, k, b, w, ctx->d_log, ctx->d_exp,
                         d_status);
      P2H_CHECK(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_gj_swap, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, aug, m, k, b, w.perm);
    P2H_CHECK(ctx, hipGetLastError());
    const long gather = std::max<long>((long)(m - b) * b, b);
    hipLaunchKernelGGL(k_gj_gather, dim3((unsigned)((gather + 255) / 256)), dim3(256), 0, s, aug, m, k, b, w);
    P2H_CHECK(ctx, hipGetLastError());
    GfProblem pn;  // P' = B^-1 P
    pn.srcs = w.psrc;
    pn.dsts = w.pdst;
    pn.coeff = w.binv;
    pn.nsrc = pn.ndst = b;
    pn.words = (size_t)words;
    pn.accumulate = false;
    pn.align_words = align;
    pn.kernel = 1;
    pn.tag = "gauss_jordan_panel";
    if ((rc = run_gf(ctx, s, pn))) return rc;
    if (m > b) {
      GfProblem up;  // row_r ^= M[r, K] P'
      up.srcs = w.usrc;
      up.dsts = w.udst;
      up.coeff = w.coef;
      up.nsrc = b;
      up.ndst = m - b;
      up.words = (size_t)words;
      up.accumulate = true;
      up.align_words = align;
      up.kernel = 1;
      up.tag = "gauss_jordan_update";
      if ((rc = run_gf(ctx, s, up))) return rc;
    }
    P2H_CHECK(ctx, hipMemcpy2DAsync(aug + (long)k * 2 * m + k, 4 * M, w.newp, 4 * M, (size_t)words * 2, (size_t)b,
                                    hipMemcpyDeviceToDevice, s));
  }
  return PAR2HIP_OK;
}

int32_t launch_gauss_jordan(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, int m, int *d_status) {
  if (m > kGjMaxM) return fail(ctx, PAR2HIP_EINVAL, "erasure count above 32768");
  if (m > kGjBlockedMinM) return gauss_jordan_blocked(ctx, s, d_aug, m, d_status);
  const size_t logs = (size_t)3 * m * sizeof(uint32_t);
  ScopedTiming t(ctx, "gauss_jordan", s);
  if (m <= kGjLdsMaxM) {
    const size_t shmem = logs + (size_t)m * 2 * m * sizeof(uint16_t) + 16;
    static std::once_flag once;
    std::call_once(once, [] {
      (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gauss_jordan<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    });
    hipLaunchKernelGGL(k_gauss_jordan<true>, dim3(1), dim3(kGjThreads), shmem, s, d_aug, m, ctx->d_log, ctx->d_exp,
                       d_status);
  } else {
    hipLaunchKernelGGL(k_gauss_jordan<false>, dim3(1), dim3(kGjThreads), logs + 16, s, d_aug, m, ctx->d_log,
                       ctx->d_exp, d_status);
  }
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// Consecutive recovery exponents e_i = e0 + i make A[i][j] = x_j^e0 * x_j^i
// with x_j = c_{missing_j}: a Vandermonde matrix times diag(x_j^e0).  Its
// inverse has a closed form (Lagrange basis): with P(t) = prod_k (t + x_k)
// and Q_j(t) = P(t) / (t + x_j) = sum_i q_ji t^i, Q_j(x_k) = 0 for k != j, so
//   (A^-1)[j][i] = q_ji / (Q_j(x_j) * x_j^e0).
// One workgroup: wave 0 builds P's coefficients in LDS, then thread j finds
// Q_j(x_j) = P'(x_j) (characteristic 2: (t + x_j) Q_j'(t) vanishes at x_j, and
// P' keeps the odd-degree terms), forms s_j = 1 / (Q_j(x_j) x_j^e0), and divides
// out (t + x_j), writing s_j q_ji as row j of A^-1 into the right half of the
// m x 2m workspace (the layout the Gauss-Jordan path uses).  s_j q is formed in
// registers from s_j x^k (k = 0..15): no table lookup per entry.
// The x_j are distinct powers of 2, so the matrix is never singular.
// Multiplication by a fixed x through four 16-entry nibble tables in LDS:
// x*v = T0[v & 15] ^ T1[(v >> 4) & 15] ^ T2[(v >> 8) & 15] ^ T3[v >> 12].
__device__ __forceinline__ uint32_t mul_tab(const uint16_t *__restrict__ t, uint32_t v) {
  return (uint32_t)t[v & 15u] ^ t[16u + ((v >> 4) & 15u)] ^ t[32u + ((v >> 8) & 15u)] ^ t[48u + (v >> 12)];
}

// s * v from bs[k] = s * x^k (k = 0..15): the xor of bs[k] over the set bits of v.
__device__ __forceinline__ uint32_t mul_bits(const uint32_t (&bs)[16], uint32_t v) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) r ^= (0u - ((v >> k) & 1u)) & bs[k];
  return r;
}

// Row j of the closed-form inverse from P's coefficients (poly, LDS) and x_j's
// nibble tables (tj): h = P'(x_j) = sum over odd i of p_i x_j^(i-1) by Horner
// in x_j^2; s = 1 / (h x_j^e0) (h != 0: the x_k are distinct); then the
// synthetic division of P by (t + x_j), q_{m-1} = p_m, q_{i-1} = p_i + x_j q_i,
// with s q_i written to row[i] (and to col[i] when col is given).
__device__ __forceinline__ void vand_inverse_row(const uint16_t *__restrict__ tj, const uint16_t *__restrict__ poly,
                                                 int m, uint32_t lx, uint32_t e0,
                                                 const uint16_t *__restrict__ log_tab,
                                                 const uint16_t *__restrict__ exp_tab, uint16_t *__restrict__ row,
                                                 uint16_t *__restrict__ col) {
  uint32_t h = 0;
  for (int i = (m & 1) ? m : m - 1; i >= 1; i -= 2) h = mul_tab(tj, mul_tab(tj, h)) ^ poly[i];
  const uint32_t ld = (log_tab[h] + (uint32_t)(((uint64_t)lx * e0) % kOrder)) % kOrder;
  uint32_t bs[16];
  bs[0] = exp_tab[ld ? kOrder - ld : 0];
#pragma unroll
  for (int k = 1; k < 16; k++) {
    const uint32_t t = bs[k - 1] << 1;
    bs[k] = (t & 0x10000u) ? (t ^ kPoly) : t;
  }
  uint32_t q = poly[m];  // = 1
  for (int i = m - 1; i >= 0; i--) {
    const uint16_t w = (uint16_t)mul_bits(bs, q);
    row[i] = w;
    if (col) col[i] = w;
    if (i > 0) q = poly[i] ^ mul_tab(tj, q);
  }
}

__global__ __launch_bounds__(1024) void k_vandermonde_inverse(uint16_t *__restrict__ aug, const uint32_t *__restrict__ miss_log,
                                                             uint32_t e0, int m, const uint16_t *__restrict__ log_tab,
                                                             const uint16_t *__restrict__ exp_tab) {
  extern __shared__ uint16_t sh16[];
  uint16_t *tabs = sh16;                 // m x 64: nibble tables of x_k = c_{missing_k}
  uint16_t *poly = sh16 + (size_t)m * 64;  // m + 1 coefficients of P, low to high
  const int tid = threadIdx.x;
  for (int e = tid; e < m * 64; e += blockDim.x) {
    const int k = e >> 6, iv = e & 63, i = iv >> 4, v = iv & 15;
    const uint32_t x = (uint32_t)v << (4 * i);
    tabs[e] = x ? exp_tab[log_tab[x] + miss_log[k]] : 0;
  }
  for (int i = tid; i <= m; i += blockDim.x) poly[i] = (i == 0) ? 1 : 0;
  __syncthreads();
  if (tid < 64) {
    // P(t) = prod_k (t + x_k), built by one wave: c'_i = c_{i-1} + x_k c_i
    for (int k = 0; k < m; k++) {
      const uint16_t *tk = tabs + (size_t)k * 64;
      uint32_t nv[17];  // m <= 1024: at most 17 coefficients per lane (i = tid + 64u <= m)
      int u = 0;
      for (int i = tid; i <= k + 1; i += 64, u++) {
        const uint32_t lo = i ? poly[i - 1] : 0u;
        const uint32_t hi = (i <= k) ? poly[i] : 0u;
        nv[u] = lo ^ mul_tab(tk, hi);
      }
      __builtin_amdgcn_wave_barrier();
      u = 0;
      for (int i = tid; i <= k + 1; i += 64, u++) poly[i] = (uint16_t)nv[u];
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  for (int j = tid; j < m; j += blockDim.x)
    vand_inverse_row(tabs + (size_t)j * 64, poly, m, miss_log[j], e0, log_tab, exp_tab,
                     aug + (size_t)j * 2 * m + m, nullptr);
}

int32_t launch_vandermonde_inverse(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, const uint32_t *d_miss_log,
                                   uint32_t e0, int m) {
  if (m > kVandMaxM) return fail(ctx, PAR2HIP_EINVAL, "too many erasures for the closed-form inverse");
  ScopedTiming t(ctx, "vandermonde_inverse", s);
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vandermonde_inverse),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
  });
  hipLaunchKernelGGL(k_vandermonde_inverse, dim3(1), dim3(1024), (size_t)(m * 64 + m + 1) * sizeof(uint16_t), s, d_aug,
                     d_miss_log, e0 % kOrder, m, ctx->d_log, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

__global__ void k_build_vand(uint16_t *__restrict__ vand, size_t vs, const uint32_t *__restrict__ present_log,
                             const uint32_t *__restrict__ exps, int m, int npresent,
                             const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)m * npresent) return;
  const int i = (int)(idx / npresent), j = (int)(idx % npresent);
  // V[i][j] = c_{present_j}^{e_i}  (par2_rs.rs:395-402)
  vand[(size_t)i * vs + j] = exp_tab[fold_mod65535(present_log[j] * exps[i])];
}

int32_t launch_build_vand(par2hip_ctx *ctx, hipStream_t s, uint16_t *vand, size_t vs,
                          const uint32_t *d_present_log, const uint32_t *d_exps, int m, int npresent) {
  const long total = (long)m * npresent;
  if (total == 0) return PAR2HIP_OK;
  ScopedTiming t(ctx, "build_present_matrix", s);
  hipLaunchKernelGGL(k_build_vand, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, vand, vs, d_present_log,
                     d_exps, m, npresent, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

__global__ void k_copy_inverse(uint16_t *__restrict__ coeff, size_t cs, const uint16_t *__restrict__ aug, int m,
                               int npresent) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)m * m) return;
  const int k = (int)(idx / m), i = (int)(idx % m);
  coeff[(size_t)k * cs + npresent + i] = aug[(long)k * 2 * m + m + i];
}

int32_t launch_copy_inverse(par2hip_ctx *ctx, hipStream_t s, uint16_t *coeff, size_t cs, const uint16_t *d_aug,
                            int m, int npresent) {
  const long total = (long)m * m;
  hipLaunchKernelGGL(k_copy_inverse, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, coeff, cs, d_aug, m,
                     npresent);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

// ---------------------------------------------------------------- batched --
// One workgroup per repair job: the closed-form inverse (consecutive
// exponents), then V rows and the inverse copy into the coefficient matrix.
__global__ __launch_bounds__(1024) void k_inverse_batch(const InvJob *__restrict__ jobs,
                                                       const uint16_t *__restrict__ log_tab,
                                                       const uint16_t *__restrict__ exp_tab) {
  extern __shared__ uint16_t sh16[];
  const InvJob J = jobs[blockIdx.x];
  const int m = J.m, tid = threadIdx.x;
  uint16_t *tabs = sh16;
  uint16_t *poly = sh16 + (size_t)m * 64;
  for (int e = tid; e < m * 64; e += blockDim.x) {
    const int k = e >> 6, iv = e & 63, i = iv >> 4, v = iv & 15;
    const uint32_t x = (uint32_t)v << (4 * i);
    tabs[e] = x ? exp_tab[log_tab[x] + J.miss_log[k]] : 0;
  }
  for (int i = tid; i <= m; i += blockDim.x) poly[i] = (i == 0) ? 1 : 0;
  __syncthreads();
  if (tid < 64) {
    for (int k = 0; k < m; k++) {
      const uint16_t *tk = tabs + (size_t)k * 64;
      uint32_t nv[17];
      int u = 0;
      for (int i = tid; i <= k + 1; i += 64, u++) {
        const uint32_t lo = i ? poly[i - 1] : 0u;
        const uint32_t hi = (i <= k) ? poly[i] : 0u;
        nv[u] = lo ^ mul_tab(tk, hi);
      }
      __builtin_amdgcn_wave_barrier();
      u = 0;
      for (int i = tid; i <= k + 1; i += 64, u++) poly[i] = (uint16_t)nv[u];
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  // row j of the inverse, also into the inverse columns of C
  for (int j = tid; j < m; j += blockDim.x)
    vand_inverse_row(tabs + (size_t)j * 64, poly, m, J.miss_log[j], J.e0, log_tab, exp_tab,
                     J.aug + (size_t)j * 2 * m + m, J.coeff + (size_t)j * J.cs + J.np);
  // V rows: V[i][c] = c_{present_c}^{e_i}
  for (long e = tid; e < (long)m * J.np; e += blockDim.x) {
    const int i = (int)(e / J.np), c = (int)(e % J.np);
    J.vand[(size_t)i * J.vs + c] = exp_tab[fold_mod65535(J.plog[c] * J.exps[i])];
  }
}

int32_t launch_inverse_batch(par2hip_ctx *ctx, hipStream_t s, const InvJob *d_jobs, int njobs, int max_m,
                             long max_vand) {
  (void)max_vand;
  if (njobs == 0) return PAR2HIP_OK;
  if (max_m > kVandMaxM) return fail(ctx, PAR2HIP_EINVAL, "too many erasures for the closed-form inverse");
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_inverse_batch),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
  });
  ScopedTiming t(ctx, "vandermonde_inverse", s);
  hipLaunchKernelGGL(k_inverse_batch, dim3((unsigned)njobs), dim3(1024), (size_t)(max_m * 64 + max_m + 1) * 2, s,
                     d_jobs, ctx->d_log, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

}  // namespace par2hip
