// Block reductions and the small symmetric pseudo-inverse shared by the
// Anderson-type iterations (AAR Poisson solve, Periodic-Pulay mixing).
#pragma once
#include "cyx_internal.hpp"

namespace cyx {

constexpr int AB = 256;  // threads per block of the reduction kernels

// sum of v over the block (AB threads), result valid in every thread
__device__ inline int tlin() { return threadIdx.x + blockDim.x * threadIdx.y; }
__device__ inline int nblocks() { return gridDim.x * gridDim.y * gridDim.z; }
__device__ inline int blin() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

__device__ inline double block_sum(double v, double* red) {
  const int t = tlin();
  red[t] = v;
  __syncthreads();
  for (int w = AB / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

// fixed-order sum of nd columns of part[nblk][nd] by the whole block into tot[nd]
__device__ inline void reduce_partials(const double* part, int nblk, int nd, double* tot) {
  __shared__ double tmp[AB];
  const int lanes = AB / nd;
  const int t = tlin();
  if (t < nd * lanes) {
    const int d = t / lanes, l = t % lanes;
    double s = 0;
    for (int b = l; b < nblk; b += lanes) s += part[(int64_t)b * nd + d];
    tmp[t] = s;
  }
  __syncthreads();
  if (t < nd) {
    double s = 0;
    for (int l = 0; l < lanes; ++l) s += tmp[t * lanes + l];
    tot[t] = s;
  }
  __syncthreads();
}

// symmetric M x M pseudo-inverse times vector with MATLAB pinv's tolerance
// max(size) * norm * eps, by cyclic Jacobi eigen-decomposition (one thread;
// compile-time M keeps A and V in registers)
template <int M>
__device__ void pinv_solve(double* A, const double* b, double* out) {
  double V[M * M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) V[i * M + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        tot += A[i * M + j] * A[i * M + j];
        if (i != j) off += A[i * M + j] * A[i * M + j];
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < M - 1; ++p)
#pragma unroll
      for (int q = p + 1; q < M; ++q) {
        const double apq = A[p * M + q];
        if (apq != 0.0) {
          const double theta = (A[q * M + q] - A[p * M + p]) / (2.0 * apq);
          const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
          for (int k = 0; k < M; ++k) {
            const double akp = A[k * M + p], akq = A[k * M + q];
            A[k * M + p] = c * akp - s * akq;
            A[k * M + q] = s * akp + c * akq;
          }
#pragma unroll
          for (int k = 0; k < M; ++k) {
            const double apk = A[p * M + k], aqk = A[q * M + k];
            A[p * M + k] = c * apk - s * aqk;
            A[q * M + k] = s * apk + c * aqk;
          }
#pragma unroll
          for (int k = 0; k < M; ++k) {
            const double vkp = V[k * M + p], vkq = V[k * M + q];
            V[k * M + p] = c * vkp - s * vkq;
            V[k * M + q] = s * vkp + c * vkq;
          }
        }
      }
  }
  double smax = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) smax = fmax(smax, fabs(A[i * M + i]));
  const double tol = M * smax * 2.220446049250313e-16;
#pragma unroll
  for (int i = 0; i < M; ++i) out[i] = 0;
#pragma unroll
  for (int e = 0; e < M; ++e) {
    const double lam = A[e * M + e];
    if (fabs(lam) > tol) {
      double vb = 0;
#pragma unroll
      for (int k = 0; k < M; ++k) vb += V[k * M + e] * b[k];
      vb /= lam;
#pragma unroll
      for (int k = 0; k < M; ++k) out[k] += V[k * M + e] * vb;
    }
  }
}

}  // namespace cyx
