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

// Wave-level helpers.  After the xor butterfly every lane holds the same
// bits (each step adds a commutative pair), so decisions taken on the result
// are uniform across the wave.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// orders the LDS accesses of one wavefront's lanes (no hardware barrier)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Recursive-halving wave reduction of S per-lane values: at distance D the
// lanes with bit D clear keep the first half of the values and those with it
// set keep the second half, each adding its partner's copy, so a level costs
// ceil(S/2) shuffles instead of S (K = 36 values: 38 shuffles rather than
// 216).  After the last level each lane holds the wave sum of one index
// (idx, or -1 for a padding slot).
template <int S, int D>
__device__ inline void halve_sum(const double (&a)[S], int lane, int base, int real, double& out, int& idx) {
  if constexpr (D == 0) {
    out = a[0];
    idx = real == 1 ? base : -1;
  } else {
    constexpr int H = (S + 1) / 2;
    const bool up = (lane & D) != 0;
    double nxt[H];
#pragma unroll
    for (int t = 0; t < H; ++t) {
      const double lo = a[t];
      const double hi = (H + t < S) ? a[H + t] : 0.0;
      nxt[t] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, D, 64);
    }
    halve_sum<H, D / 2>(nxt, lane, base + (up ? H : 0), up ? max(0, real - H) : min(real, H), out, idx);
  }
}

// K block-wide sums of per-thread values v[0..K) over a 1-D block of NT
// threads (K <= 64): halving wave reductions, then one barrier; thread k < K
// returns sum k (other threads return 0).  red holds (NT / 64) * K doubles.
template <int K, int NT>
__device__ inline double block_sums(const double* v, double* red) {
  static_assert(K >= 1 && K <= 64, "block_sums handles 1..64 values");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = v[k];
  double out;
  int idx;
  halve_sum<K, 32>(a, lane, 0, K, out, idx);
  if (idx >= 0) red[w * K + idx] = out;
  __syncthreads();
  double s = 0;
  if ((int)threadIdx.x < K) {
#pragma unroll
    for (int ww = 0; ww < NT / 64; ++ww) s += red[ww * K + threadIdx.x];
  }
  return s;
}

// The same sums for a compile-time column count by a 1-D block of NT
// threads: thread t sums rows t, t + NT, ... of every column (all loads of a
// thread independent), then block_sums; thread d < NDC returns column d.
template <int NDC, int NT>
__device__ inline double column_sums_all(const double* __restrict__ part, int nblk, double* red) {
  double v[NDC];
#pragma unroll
  for (int d = 0; d < NDC; ++d) v[d] = 0;
  for (int b = threadIdx.x; b < nblk; b += NT) {
#pragma unroll
    for (int d = 0; d < NDC; ++d) v[d] += part[(int64_t)d * nblk + b];
  }
  return block_sums<NDC, NT>(v, red);
}

// Fixed-order sums of nd columns of part[nd][nblk] (column-major partials)
// into tot[nd] by a 1-D block: wave w sums columns w, w + nwaves, ...; the
// caller synchronises before reading tot.
__device__ inline void column_sums(const double* __restrict__ part, int nblk, int nd, double* tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int d = w; d < nd; d += nw) {
    const double* col = part + (int64_t)d * nblk;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int b = lane;
    for (; b + 192 < nblk; b += 256) {
      a0 += col[b];
      a1 += col[b + 64];
      a2 += col[b + 128];
      a3 += col[b + 192];
    }
    for (; b < nblk; b += 64) a0 += col[b];
    const double s = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) tot[d] = s;
  }
}

// The pseudo-inverse solve of pinv_solve by the 64 lanes of one wavefront:
// parallel cyclic Jacobi with the round-robin ordering (the M(M-1)/2
// rotations of a sweep in M_even - 1 rounds of disjoint pairs, each round
// applied as J^T A J with all its rotations at once).  A, V and b live in
// LDS; lanes synchronise with wave_sync only.  M * ceil(M/2) <= 64.
template <int M>
struct PinvShared {
  static constexpr int ME = M + (M & 1);
  double A[M * M], V[M * M], b[M], vb[M], out[M];
};

template <int M>
__device__ void pinv_solve_wave(PinvShared<M>& W, const int lane) {
  constexpr int ME = PinvShared<M>::ME, NP = ME / 2;
  static_assert(M * NP <= 64, "one lane per (row, pair)");
  for (int e = lane; e < M * M; e += 64) W.V[e] = (e / M == e % M) ? 1.0 : 0.0;
  wave_sync();
  for (int sweep = 0; sweep < 60; ++sweep) {
    double a2 = 0, o2 = 0;
    for (int e = lane; e < M * M; e += 64) {
      const double v = W.A[e];
      a2 += v * v;
      if (e / M != e % M) o2 += v * v;
    }
    a2 = wave_sum(a2);
    o2 = wave_sum(o2);
    if (o2 <= 1e-32 * a2 || o2 == 0.0) break;
    for (int r = 0; r < ME - 1; ++r) {
      // lane it handles row/column k = it % M of pair i = it / M in both
      // half-steps, so it computes that pair's rotation itself from the A of
      // the round start
      const int it = lane, k = it % M, i = it / M;
      const bool act = it < M * NP;
      int pp = i == 0 ? r : (r + i) % (ME - 1);
      int qq = i == 0 ? ME - 1 : (r - i + ME - 1) % (ME - 1);
      if (pp > qq) {
        const int x = pp;
        pp = qq;
        qq = x;
      }
      double c = 1.0, sn = 0.0;
      if (act && qq < M) {
        const double apq = W.A[pp * M + qq];
        if (apq != 0.0) {
          const double theta = (W.A[qq * M + qq] - W.A[pp * M + pp]) / (2.0 * apq);
          const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(tt * tt + 1.0);
          sn = tt * c;
        }
      }
      const bool rot = act && qq < M && sn != 0.0;
      double akp = 0, akq = 0;
      if (rot) {
        akp = W.A[k * M + pp];
        akq = W.A[k * M + qq];
      }
      wave_sync();
      if (rot) {  // columns: A J
        W.A[k * M + pp] = c * akp - sn * akq;
        W.A[k * M + qq] = sn * akp + c * akq;
      }
      wave_sync();
      if (rot) {  // rows: J^T (A J), and V J
        const double apk = W.A[pp * M + k], aqk = W.A[qq * M + k];
        W.A[pp * M + k] = c * apk - sn * aqk;
        W.A[qq * M + k] = sn * apk + c * aqk;
        const double vkp = W.V[k * M + pp], vkq = W.V[k * M + qq];
        W.V[k * M + pp] = c * vkp - sn * vkq;
        W.V[k * M + qq] = sn * vkp + c * vkq;
      }
      wave_sync();
    }
  }
  double smax = 0;
  for (int i = 0; i < M; ++i) smax = fmax(smax, fabs(W.A[i * M + i]));
  const double tol = M * smax * 2.220446049250313e-16;
  if (lane < M) {  // vb_e = (V_e . b) / lambda_e on the kept eigenpairs
    const double lam = W.A[lane * M + lane];
    double vb = 0;
    if (fabs(lam) > tol) {
      for (int k = 0; k < M; ++k) vb += W.V[k * M + lane] * W.b[k];
      vb /= lam;
    }
    W.vb[lane] = vb;
  }
  wave_sync();
  if (lane < M) {
    double o = 0;
    for (int e = 0; e < M; ++e) o += W.V[lane * M + e] * W.vb[e];
    W.out[lane] = o;
  }
  wave_sync();
}

// pinv(A) b for a symmetric positive semi-definite A (M x M in W.A, b in
// W.b) when A is certified full rank under pinv's tolerance: Cholesky A = L
// L^T, then lambda_min(A) = 1 / ||L^{-1}||_2^2 >= 1 / ||L^{-1}||_F^2 > 100 M
// trace(A) eps >= 100 M lambda_max eps, so pinv keeps every eigenvalue and
// pinv(A) b = A^{-1} b = L^{-T} L^{-1} b.  Returns false (W.out untouched,
// W.A intact) when the certificate fails; the caller then takes
// pinv_solve_wave.  Every lane of the wavefront computes the factor, its
// inverse and the two triangular solves redundantly in registers (identical
// operations, so identical results and uniform branches); the only
// cross-lane step is the Frobenius-norm wave_sum, where lane e holds element
// e of L^{-1} in row-major order.  (An LDS form with three wavefront syncs
// per elimination step took 11.8 us per Anderson finalise at m = 7, this one
// 9.2 us.)
template <int M>
__device__ bool chol_solve_regs(PinvShared<M>& W, const int lane) {
  static_assert(M * M <= 64, "one lane per element of the inverse factor");
  double V[M][M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) V[i][j] = j <= i ? W.A[i * M + j] : 0.0;
  double tr = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) tr += W.A[i * M + i];
#pragma unroll
  for (int k = 0; k < M; ++k) {  // right-looking Cholesky, L in the lower triangle of V
    const double d = V[k][k];
    if (!(d > 0.0)) return false;
    const double lk = sqrt(d), inv = 1.0 / lk;
#pragma unroll
    for (int i = k + 1; i < M; ++i) V[i][k] = V[i][k] * inv;
    V[k][k] = lk;
#pragma unroll
    for (int j = k + 1; j < M; ++j)
#pragma unroll
      for (int i = j; i < M; ++i) V[i][j] -= V[i][k] * V[j][k];
  }
  double Li[M][M];  // L^{-1} by forward substitution, column by column
#pragma unroll
  for (int c = 0; c < M; ++c)
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double s = i == c ? 1.0 : 0.0;
      if (i >= c) {
#pragma unroll
        for (int kk = c; kk < i; ++kk) s -= V[i][kk] * Li[kk][c];
        s /= V[i][i];
      }
      Li[i][c] = s;
    }
  double mine = 0;  // lane e holds element e of L^{-1} (row-major)
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int c = 0; c < M; ++c)
      if (lane == i * M + c) mine = Li[i][c];
  double f2 = 0;
  f2 += mine * mine;
  f2 = wave_sum(f2);
  if (!(f2 > 0.0) || !(1.0 / f2 > 100.0 * M * tr * 2.220446049250313e-16)) return false;
  double bb[M], y[M];
#pragma unroll
  for (int c = 0; c < M; ++c) bb[c] = W.b[c];
#pragma unroll
  for (int l = 0; l < M; ++l) {  // y = L^{-1} b
    double t = 0;
#pragma unroll
    for (int c = 0; c <= l; ++c) t += Li[l][c] * bb[c];
    y[l] = t;
  }
#pragma unroll
  for (int l = 0; l < M; ++l) {  // x = L^{-T} y
    double t = 0;
#pragma unroll
    for (int i = l; i < M; ++i) t += Li[i][l] * y[i];
    if (lane == l) W.out[l] = t;
  }
  wave_sync();
  return true;
}

}  // namespace cyx
