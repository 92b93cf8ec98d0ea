// Gaussian likelihood algebra on the device (likelihood.py:9-61,
// sparse.py:295-389), for the block-diagonal covariances the path produces:
// the reference's sparse format [P][P][n_ell] holds, for each ell, a P x P
// block (entry (p, q, l) couples C_p(ell_l) and C_q(ell_l)).
//
// k_block_inv:  per ell, the inverse of the P x P block by Gauss-Jordan
//               elimination with partial pivoting in LDS (FP64; any
//               invertible block, as sparse.py:295-315 inverts) and its
//               pivots, whose log-magnitudes sum to log |det| (sparse.py:322-366).
// k_chi2:       per (ell, group of 32 cosmologies), r = data - mu and
//               s = C_l^-1 r with the precision block staged in LDS; writes
//               s (the derivative of the log-likelihood with respect to mu,
//               the adjoint's cotangent) and r^T s per (cosmology, ell).
// k_chi2_sum:   per cosmology, -1/2 sum_l r^T s in a fixed order plus a
//               constant (deterministic).
#include "jc_kernels.hpp"

namespace jc {

constexpr int LK_MAXP = 64;  // pairs per block (10 tracers give 55)

__global__ __launch_bounds__(256) void k_block_inv(const double* __restrict__ cov, int P, int n_ell,
                                                    double* __restrict__ inv, double* __restrict__ piv) {
  extern __shared__ double A[];  // [P][2P]: the block, then its inverse
  __shared__ int s_row, s_bad;
  const int l = blockIdx.x;
  const int W2 = 2 * P;
  // The first pass eliminates on the diagonal (the order that suits the
  // symmetric positive definite covariances, whose badly scaled rows row
  // exchanges would mix); a block with a non-positive pivot there is
  // started again with partial pivoting, which inverts any invertible block.
  for (int pivoting = 0; pivoting < 2; ++pivoting) {
    for (int e = threadIdx.x; e < P * W2; e += blockDim.x) {
      const int p = e / W2, q = e - p * W2;
      A[e] = q < P ? cov[((size_t)p * P + q) * n_ell + l] : (q - P == p ? 1.0 : 0.0);
    }
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    for (int k = 0; k < P; ++k) {
      if (pivoting) {
        // the row at or below k with the largest |A[i][k]|
        if (threadIdx.x == 0) {
          int r = k;
          double best = fabs(A[k * W2 + k]);
          for (int i = k + 1; i < P; ++i) {
            const double v = fabs(A[i * W2 + k]);
            if (v > best) {
              best = v;
              r = i;
            }
          }
          s_row = r;
        }
        __syncthreads();
        const int r = s_row;
        if (r != k) {
          for (int q = threadIdx.x; q < W2; q += blockDim.x) {
            const double t = A[k * W2 + q];
            A[k * W2 + q] = A[r * W2 + q];
            A[r * W2 + q] = t;
          }
          __syncthreads();
        }
      }
      const double pk = A[k * W2 + k];
      if (threadIdx.x == 0) {
        piv[(size_t)l * P + k] = pk;
        if (!pivoting && !(pk > 0.0)) s_bad = 1;
      }
      const double ip = 1.0 / pk;
      __syncthreads();
      // row k scaled, then every other row i -= A[i][k] * row k
      for (int q = threadIdx.x; q < W2; q += blockDim.x) A[k * W2 + q] *= ip;
      __syncthreads();
      for (int e = threadIdx.x; e < P * W2; e += blockDim.x) {
        const int i = e / W2, q = e - i * W2;
        if (i != k && q != k) A[e] -= A[i * W2 + k] * A[k * W2 + q];
      }
      __syncthreads();
      for (int i = threadIdx.x; i < P; i += blockDim.x)
        if (i != k) A[i * W2 + k] = 0.0;
      __syncthreads();
    }
    if (!s_bad) break;  // uniform: read after the last barrier
    __syncthreads();    // every thread has read s_bad before the reload resets it
  }
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) {
    const int p = e / P, q = e - p * P;
    inv[((size_t)p * P + q) * n_ell + l] = A[p * W2 + P + q];
  }
}

// a block per (ell, group of LK_CB cosmologies): the precision block is
// staged once and each wave takes every 4th cosmology of the group; lane p
// owns pair p
constexpr int LK_CB = 32;

__global__ __launch_bounds__(256) void k_chi2(const float* __restrict__ mu, int B, int P, int n_ell,
                                               const double* __restrict__ data, const double* __restrict__ prec,
                                               float* __restrict__ s_out, double* __restrict__ part) {
  __shared__ double Pl[LK_MAXP * LK_MAXP];  // column-major: lane p reads Pl[q][p], no bank conflicts
  __shared__ double rr[4][LK_MAXP];
  const int l = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) {
    const int p = e / P, q = e - p * P;
    Pl[q * LK_MAXP + p] = prec[((size_t)p * P + q) * n_ell + l];
  }
  const double dl = lane < P ? data[(size_t)lane * n_ell + l] : 0.0;
  const int b0 = blockIdx.y * LK_CB;
  for (int k = 0; k < LK_CB; k += 4) {
    const int b = b0 + k + wave;  // uniform over the block's loop: every wave runs every pass
    const bool on = b < B && lane < P;
    const double r = on ? dl - (double)mu[((size_t)b * P + lane) * n_ell + l] : 0.0;
    __syncthreads();  // the previous pass has read rr (and, first, Pl is staged)
    rr[wave][lane] = r;
    __syncthreads();
    if (b < B) {
      double s = 0.0;
      if (on)
        for (int q = 0; q < P; ++q) s += Pl[q * LK_MAXP + lane] * rr[wave][q];
      if (on && s_out) s_out[((size_t)b * P + lane) * n_ell + l] = (float)s;
      double v = r * s;
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) part[(size_t)b * n_ell + l] = v;
    }
  }
}

// one wave per cosmology: lane j adds ell j, j + 64, ..., then a butterfly
// over the wave (a fixed order)
__global__ __launch_bounds__(64) void k_chi2_sum(const double* __restrict__ part, int n_ell, double c0,
                                                  double* __restrict__ out) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0.0;
  for (int l = lane; l < n_ell; l += 64) s += part[(size_t)b * n_ell + l];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0) out[b] = -0.5 * s + c0;
}

hipError_t launch_block_inv(const double* cov, int P, int n_ell, double* inv, double* piv, hipStream_t s) {
  if (P < 1 || P > LK_MAXP) return hipErrorInvalidValue;
  const size_t sm = (size_t)P * 2 * P * sizeof(double);
  hipLaunchKernelGGL(k_block_inv, dim3(n_ell), dim3(256), sm, s, cov, P, n_ell, inv, piv);
  return hipGetLastError();
}

hipError_t launch_loglike(const float* mu, int B, int P, int n_ell, const double* data, const double* prec,
                          double c0, float* s_out, double* part, double* loglike, hipStream_t s) {
  if (P < 1 || P > LK_MAXP) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_chi2, dim3(n_ell, (B + LK_CB - 1) / LK_CB), dim3(256), 0, s, mu, B, P, n_ell, data, prec,
                     s_out, part);
  hipLaunchKernelGGL(k_chi2_sum, dim3(B), dim3(64), 0, s, part, n_ell, c0, loglike);
  return hipGetLastError();
}

}  // namespace jc
