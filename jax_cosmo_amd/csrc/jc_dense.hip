// Dense Gaussian likelihood algebra on the device (likelihood.py:43-61, the
// covariance as one N x N matrix, N = pairs x ell: 5500 for 10 tracers and
// 100 ell).  FP64 throughout, row-major N x N (numpy's layout; the matrix is
// symmetric, so rows and columns are interchangeable on input).
//
// Cholesky C = L L^T, right-looking in 64-column panels:
//   k_potf2       the panel's 64 x 64 diagonal block, unblocked, in LDS; the
//                 logarithms of its diagonal (log det C = 2 sum log L_ii) and
//                 a flag for a non-positive pivot;
//   k_trsm_panel  the panel below it, L21 = A21 L11^-T, one row per thread;
//   k_syrk        the trailing lower triangle, A22 -= L21 L21^T, 64 x 64 tiles
//                 (4 x 4 per thread) from LDS.  FP64 FMA issues at the rate
//                 of the FP64 matrix cores on MI355X (78.6 TFLOP/s both), so
//                 the update stays on the VALU.
// Triangular solves for B right-hand sides stored as rows ([B][N]), in
// 64-row blocks: k_trsv_diag solves a block's triangle, k_solve_update
// removes it from the remaining rows (z = L^-1 r forwards, y = L^-T z
// backwards).
// The likelihood of a batch of means: r = data - mu, z = L^-1 r,
// -1/2 |z|^2 + c0 per mean (fixed-order sums), and its derivative with
// respect to the mean, C^-1 r = L^-T z, the adjoint's cotangent.
#include <algorithm>

#include "jc_kernels.hpp"

namespace jc {

constexpr int DB = 64;  // panel width and tile size

__global__ __launch_bounds__(256) void k_potf2(double* __restrict__ A, int64_t N, int k0, double* __restrict__ dlog,
                                               int* __restrict__ bad) {
  __shared__ double S[DB][DB + 1];
  const int nb = (int)min((int64_t)DB, N - k0);
  for (int e = threadIdx.x; e < DB * DB; e += blockDim.x) {
    const int i = e / DB, j = e - i * DB;
    S[i][j] = (i < nb && j < nb && j <= i) ? A[(size_t)(k0 + i) * N + k0 + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (threadIdx.x == 0) {
      double d = S[j][j];
      if (!(d > 0.0)) {
        *bad = 1;
        d = 1.0;  // keeps the rest finite; the caller reports the failure
      }
      d = sqrt(d);
      S[j][j] = d;
      dlog[k0 + j] = log(d);
    }
    __syncthreads();
    const double id = 1.0 / S[j][j];
    for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) S[i][j] *= id;
    __syncthreads();
    // trailing lower triangle of the block: (i, c) with j < c <= i
    const int m = nb - j - 1;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = j + 1 + e / m, c = j + 1 + e % m;
      if (c <= i) S[i][c] -= S[i][j] * S[c][j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e - i * nb;
    if (j <= i) A[(size_t)(k0 + i) * N + k0 + j] = S[i][j];
  }
}

// rows i >= k0 + 64: x = a L11^-T, i.e. x_j = (a_j - sum_{t<j} x_t L_jt) / L_jj
__global__ __launch_bounds__(256) void k_trsm_panel(double* __restrict__ A, int64_t N, int k0) {
  __shared__ double L11[DB][DB];
  for (int e = threadIdx.x; e < DB * DB; e += blockDim.x) {
    const int i = e / DB, j = e - i * DB;
    L11[i][j] = (j <= i) ? A[(size_t)(k0 + i) * N + k0 + j] : 0.0;
  }
  __syncthreads();
  const int64_t i = (int64_t)k0 + DB + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double* row = A + (size_t)i * N + k0;
  double x[DB];
#pragma unroll
  for (int t = 0; t < DB; ++t) x[t] = row[t];
#pragma unroll
  for (int j = 0; j < DB; ++j) {
    double s = x[j];
#pragma unroll
    for (int t = 0; t < j; ++t) s -= x[t] * L11[j][t];
    x[j] = s / L11[j][j];
  }
#pragma unroll
  for (int t = 0; t < DB; ++t) row[t] = x[t];
}

// trailing update of the lower triangle from row/column r0 = k0 + 64: tile
// (ti, tj), tj <= ti, of 64 x 64; thread (tr, tc) owns rows 4 tr .. 4 tr + 3
// and columns 4 tc .. 4 tc + 3 of the tile
__global__ __launch_bounds__(256) void k_syrk(double* __restrict__ A, int64_t N, int k0) {
  __shared__ double Pi[DB][DB + 2];  // [t][row of tile i]
  __shared__ double Pj[DB][DB + 2];  // [t][row of tile j]
  const int64_t r0 = (int64_t)k0 + DB;
  const int64_t idx = blockIdx.x;
  int64_t ti = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > idx) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= idx) ++ti;
  const int64_t tj = idx - ti * (ti + 1) / 2;
  const int64_t ri = r0 + ti * DB, rj = r0 + tj * DB;
  for (int e = threadIdx.x; e < DB * DB; e += blockDim.x) {
    const int r = e / DB, t = e - r * DB;
    Pi[t][r] = (ri + r < N) ? A[(size_t)(ri + r) * N + k0 + t] : 0.0;
    Pj[t][r] = (rj + r < N) ? A[(size_t)(rj + r) * N + k0 + t] : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
#pragma unroll 8
  for (int t = 0; t < DB; ++t) {
    const double2 a01 = *reinterpret_cast<const double2*>(&Pi[t][tr * 4]);
    const double2 a23 = *reinterpret_cast<const double2*>(&Pi[t][tr * 4 + 2]);
    const double2 b01 = *reinterpret_cast<const double2*>(&Pj[t][tc * 4]);
    const double2 b23 = *reinterpret_cast<const double2*>(&Pj[t][tc * 4 + 2]);
    const double av[4] = {a01.x, a01.y, a23.x, a23.y}, bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t row = ri + tr * 4 + a;
    if (row >= N) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t col = rj + tc * 4 + b;
      if (col <= row && col < N) A[(size_t)row * N + col] -= acc[a][b];
    }
  }
}

// the strict upper triangle of the factor is zeroed at the end
__global__ void k_zero_upper(double* __restrict__ A, int64_t N) {
  const int64_t row = blockIdx.y;
  for (int64_t c = row + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x)
    A[(size_t)row * N + c] = 0.0;
}

// Triangular solves for B right-hand sides stored as rows of Z ([B][N]), in
// 64-row blocks.  For each block: solve its triangle for every right-hand
// side (k_trsv_diag), then remove the solved block from the rows still to
// come (k_solve_update).
//   forward (z = L^-1 r):    rows after the block, z_i -= sum_t L[i][k0+t] z_{k0+t}
//   transpose (y = L^-T z):  rows before the block, y_i -= sum_t L[k0+t][i] y_{k0+t}
// k_trsv_diag: the block's 64 x 64 triangle staged in LDS once per four
// right-hand sides; one wave per right-hand side, lane i holding row k0 + i,
// the solved value of step j passed to the other lanes by a shuffle (no
// barriers inside the solve).
__global__ __launch_bounds__(256) void k_trsv_diag(const double* __restrict__ L, int64_t N, int k0, int transpose,
                                                   int64_t B, double* __restrict__ Z) {
  __shared__ double T[DB][DB + 1];
  const int nb = (int)min((int64_t)DB, N - k0);
  for (int e = threadIdx.x; e < DB * DB; e += blockDim.x) {
    const int r = e / DB, c = e - r * DB;
    T[r][c] = (r < nb && c < nb && c <= r) ? L[(size_t)(k0 + r) * N + k0 + c] : 0.0;
  }
  __syncthreads();
  const int i = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= B) return;  // after the block's only barrier
  double* z = Z + (size_t)b * N;
  double s = i < nb ? z[k0 + i] : 0.0;  // rows past nb stay zero (their T entries are zero)
  if (!transpose) {
    for (int j = 0; j < nb; ++j) {
      const double zj = __shfl(s, j, 64) / T[j][j];
      if (i == j) s = zj;
      else if (i > j) s -= T[i][j] * zj;
    }
  } else {
    for (int j = nb - 1; j >= 0; --j) {
      const double zj = __shfl(s, j, 64) / T[j][j];
      if (i == j) s = zj;
      else if (i < j) s -= T[j][i] * zj;
    }
  }
  if (i < nb) z[k0 + i] = s;
}

// k_solve_update: rows r of the tile (64 per block, grid x) and 16
// right-hand sides (grid y); thread (row r = t % 64, group g = t / 64)
// updates 4 right-hand sides from the 64 x 64 tile of L and the solved
// values in LDS.  (Measured: 64 right-hand sides per block with 4 x 4 per
// thread ran 25 us per step against 19 for this form, fewer blocks in
// flight at the small trailing sizes.)
constexpr int SU_R = 16;  // right-hand sides per update block

__global__ __launch_bounds__(256) void k_solve_update(const double* __restrict__ L, int64_t N, int k0, int transpose,
                                                      int64_t B, double* __restrict__ Z) {
  __shared__ double Lt[DB][DB + 1];   // [t][r]: the coupling of row r to solved row k0 + t
  __shared__ double zs[SU_R][DB];     // solved values of the block's right-hand sides
  const int nb = (int)min((int64_t)DB, N - k0);
  // the rows this block updates: after the block (forward) or before it (transpose)
  const int64_t row0 = transpose ? (int64_t)blockIdx.x * DB : (int64_t)k0 + DB + (int64_t)blockIdx.x * DB;
  const int64_t row_end = transpose ? (int64_t)k0 : N;
  const int64_t b0 = (int64_t)blockIdx.y * SU_R;
  const int t = threadIdx.x;
  for (int e = t; e < DB * DB; e += 256) {
    const int x = e / DB, y = e - x * DB;
    double v = 0.0;
    if (!transpose) {
      // x = row offset, y = t: L[row0 + x][k0 + y] (coalesced in y)
      if (row0 + x < row_end && y < nb) v = L[(size_t)(row0 + x) * N + k0 + y];
      Lt[y][x] = v;
    } else {
      // x = t, y = row offset: L[k0 + x][row0 + y] (coalesced in y)
      if (x < nb && row0 + y < row_end) v = L[(size_t)(k0 + x) * N + row0 + y];
      Lt[x][y] = v;
    }
  }
  for (int e = t; e < SU_R * DB; e += 256) {
    const int q = e / DB, c = e - q * DB;
    zs[q][c] = (b0 + q < B && c < nb) ? Z[(size_t)(b0 + q) * N + k0 + c] : 0.0;
  }
  __syncthreads();
  const int r = t & (DB - 1), g = t >> 6;  // 4 groups of 4 right-hand sides
  const int64_t row = row0 + r;
  if (row >= row_end) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < nb; ++c) {
    const double l = Lt[c][r];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += l * zs[g * 4 + q][c];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t bb = b0 + g * 4 + q;
    if (bb < B) Z[(size_t)bb * N + row] -= acc[q];
  }
}

// r = data - mu (FP64), [B][N]
__global__ void k_resid(const float* __restrict__ mu, const double* __restrict__ data, int64_t N, int64_t B,
                        double* __restrict__ r) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * B) return;
  r[e] = data[e % N] - (double)mu[e];
}

// out[b] = -1/2 sum_i z[b][i]^2 + c0: 256 threads per right-hand side, fixed order
__global__ __launch_bounds__(256) void k_sq_sum(const double* __restrict__ Z, int64_t N, double c0,
                                                double* __restrict__ out) {
  __shared__ double ws[4];
  const int64_t b = blockIdx.x;
  const double* z = Z + (size_t)b * N;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += 256) s += z[i] * z[i];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[b] = -0.5 * (((ws[0] + ws[1]) + ws[2]) + ws[3]) + c0;
}

__global__ void k_to_float(const double* __restrict__ x, int64_t n, float* __restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = (float)x[e];
}

hipError_t launch_dense_cholesky(double* A, int64_t N, double* dlog, int* bad, hipStream_t s) {
  if (N < 1) return hipErrorInvalidValue;
  for (int64_t k0 = 0; k0 < N; k0 += DB) {
    hipLaunchKernelGGL(k_potf2, dim3(1), dim3(256), 0, s, A, N, (int)k0, dlog, bad);
    const int64_t rest = N - k0 - DB;
    if (rest > 0) {
      hipLaunchKernelGGL(k_trsm_panel, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s, A, N, (int)k0);
      const int64_t T = (rest + DB - 1) / DB;
      hipLaunchKernelGGL(k_syrk, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, A, N, (int)k0);
    }
  }
  hipLaunchKernelGGL(k_zero_upper, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, 64)),
                                         (unsigned)N),
                     dim3(256), 0, s, A, N);
  return hipGetLastError();
}

hipError_t launch_dense_solve(const double* L, int64_t N, double* Z, int64_t B, int transpose, hipStream_t s) {
  if (N < 1 || B < 1) return hipErrorInvalidValue;
  const unsigned gy = (unsigned)((B + SU_R - 1) / SU_R);
  const int64_t last = (N - 1) / DB * DB;
  for (int64_t step = 0; step <= last; step += DB) {
    const int64_t k0 = transpose ? last - step : step;
    hipLaunchKernelGGL(k_trsv_diag, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, s, L, N, (int)k0, transpose, B, Z);
    const int64_t rows = transpose ? k0 : N - k0 - DB;
    if (rows > 0)
      hipLaunchKernelGGL(k_solve_update, dim3((unsigned)((rows + DB - 1) / DB), gy), dim3(256), 0, s, L, N, (int)k0,
                         transpose, B, Z);
  }
  return hipGetLastError();
}

hipError_t launch_dense_loglike(const float* mu, int64_t B, int64_t N, const double* data, const double* L, double c0,
                                double* Z, double* loglike, float* dmu, hipStream_t s) {
  const int64_t n = B * N;
  hipLaunchKernelGGL(k_resid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mu, data, N, B, Z);
  hipError_t e = launch_dense_solve(L, N, Z, B, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sq_sum, dim3((unsigned)B), dim3(256), 0, s, Z, N, c0, loglike);
  if (dmu) {
    if ((e = launch_dense_solve(L, N, Z, B, 1, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Z, n, dmu);
  }
  return hipGetLastError();
}

}  // namespace jc
