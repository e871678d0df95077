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
constexpr int kGjMaxM = 4096;
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

int32_t launch_gauss_jordan(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, int m, int *d_status) {
  if (m > kGjMaxM) return fail(ctx, PAR2HIP_EINVAL, "erasure count above the device Gauss-Jordan limit (4096)");
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
// One workgroup: wave 0 builds P's coefficients in LDS, then thread j divides
// out (t + x_j), evaluates Q_j(x_j) by Horner and writes row j of A^-1 into the
// right half of the m x 2m workspace (the layout the Gauss-Jordan path uses).
// The x_j are distinct powers of 2, so the matrix is never singular.
// Multiplication by a fixed x through four 16-entry nibble tables in LDS:
// x*v = T0[v & 15] ^ T1[(v >> 4) & 15] ^ T2[(v >> 8) & 15] ^ T3[v >> 12].
__device__ __forceinline__ uint32_t mul_tab(const uint16_t *__restrict__ t, uint32_t v) {
  return (uint32_t)t[v & 15u] ^ t[16u + ((v >> 4) & 15u)] ^ t[32u + ((v >> 8) & 15u)] ^ t[48u + (v >> 12)];
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
      uint32_t nv[9];  // m <= 512: at most 9 coefficients per lane (i = tid + 64u <= m)
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
  for (int j = tid; j < m; j += blockDim.x) {
    const uint16_t *tj = tabs + (size_t)j * 64;
    // synthetic division of P by (t + x_j): q_{m-1} = p_m, q_{i-1} = p_i + x_j q_i;
    // Horner for Q_j(x_j) runs alongside
    uint16_t *row = aug + (size_t)j * 2 * m + m;
    uint32_t q = poly[m];  // = 1
    row[m - 1] = (uint16_t)q;
    uint32_t h = q;
    for (int i = m - 1; i >= 1; i--) {
      q = poly[i] ^ mul_tab(tj, q);
      row[i - 1] = (uint16_t)q;
      h = mul_tab(tj, h) ^ q;
    }
    // h = Q_j(x_j) != 0; scale the row by 1 / (h * x_j^e0)
    const uint32_t lx = miss_log[j];
    const uint32_t ld = (log_tab[h] + (uint32_t)(((uint64_t)lx * e0) % kOrder)) % kOrder;
    const uint32_t linv = kOrder - ld;
    for (int i = 0; i < m; i++) {
      const uint32_t v = row[i];
      row[i] = v ? exp_tab[log_tab[v] + linv] : 0;
    }
  }
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
      uint32_t nv[9];
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
  for (int j = tid; j < m; j += blockDim.x) {
    const uint16_t *tj = tabs + (size_t)j * 64;
    uint16_t *row = J.aug + (size_t)j * 2 * m + m;
    uint32_t q = poly[m];
    row[m - 1] = (uint16_t)q;
    uint32_t h = q;
    for (int i = m - 1; i >= 1; i--) {
      q = poly[i] ^ mul_tab(tj, q);
      row[i - 1] = (uint16_t)q;
      h = mul_tab(tj, h) ^ q;
    }
    const uint32_t lx = J.miss_log[j];
    const uint32_t ld = (log_tab[h] + (uint32_t)(((uint64_t)lx * J.e0) % kOrder)) % kOrder;
    const uint32_t linv = kOrder - ld;
    for (int i = 0; i < m; i++) {
      const uint32_t v = row[i];
      const uint16_t w = v ? exp_tab[log_tab[v] + linv] : 0;
      row[i] = w;
      J.coeff[(size_t)j * J.cs + J.np + i] = w;  // inverse columns of C
    }
  }
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
