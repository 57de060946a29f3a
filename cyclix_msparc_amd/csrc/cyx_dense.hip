// Householder reduction of the Rayleigh-Ritz matrix to real tridiagonal form
// in one workgroup (the eigensolve of src/eigSolver.m:110, after the pencil
// has been reduced to standard form).  rocSOLVER's heevd reduces with one
// small kernel per column; at n = 256 that is ~5 ms of launch latency.  Here
// the whole reduction is one launch: the trailing matrix (1 MB at n = 256)
// stays in L2, the reflector and the product vector in LDS.
//
// Lower variant with LAPACK's storage (zhetd2 / dsytd2, uplo = 'L'):
//   for k = 0 .. n-2:  x = A[k+1:n, k],  H_k = I - tau_k v_k v_k^H with
//   v_k = (1, A[k+2:n, k]) so that H_k^H x = beta e_1 (beta real, the
//   reflector of zlarfg), then A22 := H_k^H A22 H_k as
//   A22 -= v w^H + w v^H,  w = tau A22 v - (tau/2) (w^H v) v.
//   d[k] = A[k][k], e[k] = beta, tau[k] = tau_k;  A = Q T Q^H with
//   Q = H_0 H_1 ... H_{n-2}, applied afterwards by rocsolver_{z,d}{un,or}mtr.
// Both triangles of A22 are updated (each element by its own thread, so the
// product A22 v reads columns with coalesced loads); the diagonal is kept
// real.
#include "cyx_internal.hpp"
#include "cyx_reduce.hpp"

namespace cyx {

constexpr int TRD_NT = 1024;             // threads of the reduction workgroup
constexpr int TRD_G = TRD_NT / 256;      // column groups of the matrix-vector product

__device__ __forceinline__ double re_of(double a) { return a; }
__device__ __forceinline__ double re_of(cplx<double> a) { return a.x; }
__device__ __forceinline__ double im_of(double) { return 0.0; }
__device__ __forceinline__ double im_of(cplx<double> a) { return a.y; }
template <class D>
__device__ __forceinline__ D make_d(double r, double i);
template <>
__device__ __forceinline__ double make_d<double>(double r, double) {
  return r;
}
template <>
__device__ __forceinline__ cplx<double> make_d<cplx<double>>(double r, double i) {
  return cplx<double>{r, i};
}
// a * conj(b)
__device__ __forceinline__ double mul_cj(double a, double b) { return a * b; }
__device__ __forceinline__ cplx<double> mul_cj(cplx<double> a, cplx<double> b) {
  return cplx<double>{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

template <class D>
__global__ void __launch_bounds__(TRD_NT) hetrd_kernel(D* __restrict__ A, int n, double* __restrict__ d,
                                                       double* __restrict__ e, D* __restrict__ tau) {
  extern __shared__ __align__(16) unsigned char trd_raw[];
  D* vv = reinterpret_cast<D*>(trd_raw);  // [n] reflector
  D* ww = vv + n;                         // [n] w
  D* pp = ww + n;                         // [TRD_G][n] partial products
  __shared__ double red[(TRD_NT / 64) * 2];
  __shared__ double s_beta, s_dot[2];
  __shared__ D s_tau, s_scale;
  __shared__ int s_skip;
  const int t = threadIdx.x;
  for (int k = 0; k < n - 1; ++k) {
    const int m = n - k - 1;
    D* col = A + (int64_t)k * n + (k + 1);  // x = col[0 .. m)
    // ---- reflector (zlarfg): xnorm^2 of col[1 .. m)
    double q[1] = {0.0};
    for (int i = 1 + t; i < m; i += TRD_NT) q[0] += absq(col[i]);
    const double xn2 = block_sums<1, TRD_NT>(q, red);
    if (t == 0) {
      const D alpha = col[0];
      const double ar = re_of(alpha), ai = im_of(alpha);
      if (xn2 == 0.0 && ai == 0.0) {
        s_skip = 1;
        s_tau = make_d<D>(0.0, 0.0);
        s_beta = ar;
      } else {
        const double beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
        s_tau = make_d<D>((beta - ar) / beta, -ai / beta);
        // 1 / (alpha - beta)
        const double zr = ar - beta, zi = ai, den = zr * zr + zi * zi;
        s_scale = make_d<D>(zr / den, -zi / den);
        s_beta = beta;
        s_skip = 0;
      }
      d[k] = re_of(A[(int64_t)k * n + k]);
      e[k] = s_beta;
      tau[k] = s_tau;
    }
    __syncthreads();
    if (!s_skip) {
      // ---- v = (1, col[1..m) / (alpha - beta)), stored back as LAPACK does
      for (int i = t; i < m; i += TRD_NT) {
        const D vi = i == 0 ? make_d<D>(1.0, 0.0) : cmul(s_scale, col[i]);
        vv[i] = vi;
        if (i > 0) col[i] = vi;
      }
      __syncthreads();
      // ---- A22 v: rows r (+256), columns j = g (+TRD_G); A22[i][j] = A[(k+1+j) n + k+1+i]
      {
        const int r = t & 255, g = t >> 8;
        const D* a22 = A + (int64_t)(k + 1) * n + (k + 1);
        for (int i = r; i < m; i += 256) {
          D acc = make_d<D>(0.0, 0.0);
          for (int j = g; j < m; j += TRD_G) acc = cfma(a22[(int64_t)j * n + i], vv[j], acc);
          pp[g * n + i] = acc;
        }
      }
      __syncthreads();
      // ---- w = tau A22 v and w^H v
      double dv[2] = {0.0, 0.0};
      for (int i = t; i < m; i += TRD_NT) {
        D s = pp[i];
#pragma unroll
        for (int g = 1; g < TRD_G; ++g) s = add(s, pp[g * n + i]);
        const D wi = cmul(s_tau, s);
        ww[i] = wi;
        const D c = mul_cj(vv[i], wi);  // conj(w_i) v_i
        dv[0] += re_of(c);
        dv[1] += im_of(c);
      }
      const double dsum = block_sums<2, TRD_NT>(dv, red);
      if (t < 2) s_dot[t] = dsum;
      __syncthreads();
      // ---- w += (-tau/2) (w^H v) v
      {
        const D a2 = rmul(-0.5, cmul(s_tau, make_d<D>(s_dot[0], s_dot[1])));
        for (int i = t; i < m; i += TRD_NT) ww[i] = cfma(a2, vv[i], ww[i]);
      }
      __syncthreads();
      // ---- A22 -= v w^H + w v^H
      {
        D* a22 = A + (int64_t)(k + 1) * n + (k + 1);
        for (int idx = t; idx < m * m; idx += TRD_NT) {
          const int i = idx % m, j = idx / m;
          D* a = a22 + (int64_t)j * n + i;
          D na = sub(*a, add(mul_cj(vv[i], ww[j]), mul_cj(ww[i], vv[j])));
          if (i == j) na = make_d<D>(re_of(na), 0.0);
          *a = na;
        }
      }
    }
    if (t == 0) col[0] = make_d<D>(s_beta, 0.0);
    __syncthreads();
  }
  if (t == 0) d[n - 1] = re_of(A[(int64_t)(n - 1) * n + (n - 1)]);
}

size_t hetrd_lds_bytes(int n, bool cx) { return (size_t)(2 + TRD_G) * n * (cx ? 16 : 8); }

int hetrd(hipStream_t st, bool cx, void* A, int n, double* d, double* e, void* tau) {
  const size_t shm = hetrd_lds_bytes(n, cx);
  CYX_CHECK(shm <= 64 * 1024, "tridiagonal reduction of n = %d does not fit one workgroup", n);
  if (cx)
    hipLaunchKernelGGL(hetrd_kernel<cplx<double>>, dim3(1), dim3(TRD_NT), shm, st, (cplx<double>*)A, n, d, e,
                       (cplx<double>*)tau);
  else
    hipLaunchKernelGGL(hetrd_kernel<double>, dim3(1), dim3(TRD_NT), shm, st, (double*)A, n, d, e, (double*)tau);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

}  // namespace cyx
