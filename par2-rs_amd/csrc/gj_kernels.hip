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

__global__ void k_build_aug(uint16_t *__restrict__ aug, const uint32_t *__restrict__ miss_log,
                            const uint32_t *__restrict__ exps, int m,
                            const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)m * 2 * m) return;
  const int i = (int)(idx / (2 * m)), j = (int)(idx % (2 * m));
  uint16_t v;
  if (j < m) {
    // A[i][j] = c_{missing_j}^{e_i}  (par2_rs.rs:347-359)
    v = exp_tab[((uint64_t)miss_log[j] * exps[i]) % kOrder];
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

// One workgroup.  Pivot rule as the reference: keep A[k][k] if non-zero,
// otherwise take the first row below with a non-zero entry (par2_rs.rs:362-371).
__global__ __launch_bounds__(kGjThreads) void k_gauss_jordan(uint16_t *__restrict__ aug, int m,
                                                            const uint16_t *__restrict__ log_tab,
                                                            const uint16_t *__restrict__ exp_tab,
                                                            int *__restrict__ status) {
  extern __shared__ uint32_t sh[];
  uint32_t *logrow = sh;                // 2m entries: log of pivot row, 0xFFFFFFFF for zero
  uint16_t *fac = reinterpret_cast<uint16_t *>(sh + 2 * m);  // m entries
  __shared__ int piv;
  const int tid = threadIdx.x;
  const int w = 2 * m;
  for (int k = 0; k < m; k++) {
    if (tid == 0) piv = m;
    __syncthreads();
    for (int r = k + tid; r < m; r += kGjThreads)
      if (aug[(long)r * w + k] != 0) atomicMin(&piv, r);
    __syncthreads();
    const int p = piv;
    if (p == m) {
      if (tid == 0) *status = 1;
      return;
    }
    if (p != k) {
      for (int j = tid; j < w; j += kGjThreads) {
        const uint16_t a = aug[(long)k * w + j];
        aug[(long)k * w + j] = aug[(long)p * w + j];
        aug[(long)p * w + j] = a;
      }
      __syncthreads();
    }
    // normalise the pivot row by 1/pivot and keep its logs in LDS
    const uint32_t lp = log_tab[aug[(long)k * w + k]];
    const uint32_t linv = kOrder - lp;  // log(1/p); exp is doubled so linv=65535 reads 1
    for (int j = tid; j < w; j += kGjThreads) {
      const uint16_t a = aug[(long)k * w + j];
      uint32_t l = 0xFFFFFFFFu;
      if (a) {
        const uint16_t v = exp_tab[log_tab[a] + linv];
        aug[(long)k * w + j] = v;
        l = log_tab[v];
      }
      logrow[j] = l;
    }
    for (int r = tid; r < m; r += kGjThreads) fac[r] = (r == k) ? 0 : aug[(long)r * w + k];
    __syncthreads();
    // eliminate column k from every other row (columns < k of the pivot row are zero)
    const int cols = w - k;
    const long pairs = (long)m * cols;
    for (long q = tid; q < pairs; q += kGjThreads) {
      const int r = (int)(q / cols), j = k + (int)(q % cols);
      const uint16_t f = fac[r];
      const uint32_t l = logrow[j];
      if (f && l != 0xFFFFFFFFu) aug[(long)r * w + j] ^= exp_tab[log_tab[f] + l];
    }
    __syncthreads();
  }
  if (tid == 0) *status = 0;
}

int32_t launch_gauss_jordan(par2hip_ctx *ctx, hipStream_t s, uint16_t *d_aug, int m, int *d_status) {
  if (m > kGjMaxM) return fail(ctx, PAR2HIP_EINVAL, "erasure count above the device Gauss-Jordan limit (4096)");
  const size_t shmem = (size_t)2 * m * sizeof(uint32_t) + (size_t)m * sizeof(uint16_t) + 16;
  ScopedTiming t(ctx, "gauss_jordan", s);
  hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(kGjThreads), shmem, s, d_aug, m, ctx->d_log,
                     ctx->d_exp, d_status);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

__global__ void k_fold(uint16_t *__restrict__ coeff, const uint16_t *__restrict__ aug,
                       const uint32_t *__restrict__ present_log, const uint32_t *__restrict__ exps, int m,
                       int npresent, const uint16_t *__restrict__ log_tab,
                       const uint16_t *__restrict__ exp_tab) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nsrc = npresent + m;
  if (idx >= (long)m * nsrc) return;
  const int k = (int)(idx / nsrc), j = (int)(idx % nsrc);
  const uint16_t *inv = aug + (long)k * 2 * m + m;  // row k of A^-1
  uint16_t v;
  if (j >= npresent) {
    v = inv[j - npresent];
  } else {
    // (A^-1 V)[k][j] with V[i][j] = c_{present_j}^{e_i}; log V = n_j * e_i mod 65535
    uint32_t acc = 0;
    const uint64_t lj = present_log[j];
    for (int i = 0; i < m; i++) {
      const uint16_t a = inv[i];
      if (a) acc ^= exp_tab[log_tab[a] + (uint32_t)((lj * exps[i]) % kOrder)];
    }
    v = (uint16_t)acc;
  }
  coeff[idx] = v;
}

int32_t launch_fold(par2hip_ctx *ctx, hipStream_t s, uint16_t *coeff, const uint16_t *d_aug,
                    const uint32_t *d_present_log, const uint32_t *d_exps, int m, int npresent) {
  const long total = (long)m * (npresent + m);
  ScopedTiming t(ctx, "fold_inverse", s);
  hipLaunchKernelGGL(k_fold, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, coeff, d_aug,
                     d_present_log, d_exps, m, npresent, ctx->d_log, ctx->d_exp);
  P2H_CHECK(ctx, hipGetLastError());
  return PAR2HIP_OK;
}

}  // namespace par2hip
