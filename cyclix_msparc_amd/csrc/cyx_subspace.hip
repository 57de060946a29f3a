// Subspace kernels of the Rayleigh-Ritz step (src/eigSolver.m:99-124):
// weighted Gram/projection M = A^H diag(W) B, rotation A*Q, column
// W-normalisation.  All products accumulate in FP64 (FP32 inputs are
// widened exactly, so the Gram of an FP32 block carries FP64 accuracy).
#include "cyx_internal.hpp"

namespace cyx {

constexpr int GT = 64;  // output tile
constexpr int GK = 16;  // k chunk

// Partial[kb][a][b] = sum_{n in chunk kb} conj(A[n,a]) W[n] B[n,b]  (column-major na x nb per partial)
template <class T>
__global__ void __launch_bounds__(256) gram_partial_kernel(const T* __restrict__ A, int64_t lda, int na,
                                                           const T* __restrict__ B, int64_t ldb, int nb,
                                                           const double* __restrict__ W, int64_t N, int64_t krange,
                                                           typename acc_of<T>::type* __restrict__ part) {
  using D = typename acc_of<T>::type;
  __shared__ D sa[GK][GT + 1];
  __shared__ D sb[GK][GT + 1];
  const int ta = blockIdx.x, tb = blockIdx.y, kb = blockIdx.z;
  const int a0 = ta * GT, b0 = tb * GT;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  D acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = D{};
  const int64_t k0 = (int64_t)kb * krange;
  const int64_t k1 = min(N, k0 + krange);
  // loader mapping: 256 threads load GK x GT = 1024 elements of each operand, 4 each
  for (int64_t kc = k0; kc < k1; kc += GK) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int idx = threadIdx.x + 256 * l;
      const int kk = idx % GK;
      const int col = idx / GK;
      const int64_t n = kc + kk;
      D va{}, vb{};
      if (n < k1) {
        const double w = W[n];
        if (a0 + col < na) va = widen(A[(int64_t)(a0 + col) * lda + n]);
        if (b0 + col < nb) {
          const D bb = widen(B[(int64_t)(b0 + col) * ldb + n]);
          vb = rmul(w, bb);
        }
      }
      sa[kk][col] = va;
      sb[kk][col] = vb;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      D av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) av[x] = sa[kk][tx + 16 * x];
#pragma unroll
      for (int y = 0; y < 4; ++y) bv[y] = sb[kk][ty + 16 * y];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = cjfma(av[x], bv[y], acc[x][y]);
    }
    __syncthreads();
  }
  D* out = part + (int64_t)kb * na * nb;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int a = a0 + tx + 16 * x, b = b0 + ty + 16 * y;
      if (a < na && b < nb) out[(int64_t)b * na + a] = acc[x][y];
    }
}

template <class D>
__global__ void reduce_partials_kernel(const D* __restrict__ part, int nparts, int64_t len, D* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  D s{};
  for (int k = 0; k < nparts; ++k) s = add(s, part[(int64_t)k * len + e]);
  out[e] = s;
}

// out[n,b] = sum_a A[n,a] Q[a,b]   (Q: ns x ns column-major, FP64 / complex FP64)
template <class T>
__global__ void __launch_bounds__(256) rotate_kernel(const T* __restrict__ A, const typename acc_of<T>::type* __restrict__ Q,
                                                     T* __restrict__ out, int64_t N, int64_t ld, int ns) {
  using D = typename acc_of<T>::type;
  __shared__ D sa[GK][GT + 1];  // [a][n]
  __shared__ D sq[GK][GT + 1];  // [a][b]
  const int64_t n0 = (int64_t)blockIdx.x * GT;
  const int b0 = blockIdx.y * GT;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  D acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = D{};
  for (int ac = 0; ac < ns; ac += GK) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int idx = threadIdx.x + 256 * l;
      // A tile: n fastest
      {
        const int nn = idx % GT, aa = idx / GT;
        const int64_t n = n0 + nn;
        D v{};
        if (n < N && ac + aa < ns) v = widen(A[(int64_t)(ac + aa) * ld + n]);
        sa[aa][nn] = v;
      }
      // Q tile: a fastest
      {
        const int aa = idx % GK, bb = idx / GK;
        D v{};
        if (ac + aa < ns && b0 + bb < ns) v = Q[(int64_t)(b0 + bb) * ns + ac + aa];
        sq[aa][bb] = v;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      D av[4], qv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) av[x] = sa[kk][tx + 16 * x];
#pragma unroll
      for (int y = 0; y < 4; ++y) qv[y] = sq[kk][ty + 16 * y];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = cfma(av[x], qv[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int64_t n = n0 + tx + 16 * x;
      const int b = b0 + ty + 16 * y;
      if (n < N && b < ns) out[(int64_t)b * ld + n] = narrow<T>(acc[x][y]);
    }
}

// per-column 1/sqrt(sum W |a|^2)   (src/eigSolver.m:122)
template <class T>
__global__ void __launch_bounds__(1024) colnorm_kernel(const T* __restrict__ A, const double* __restrict__ W, int64_t N,
                                                       int64_t ld, double* __restrict__ scal) {
  __shared__ double red[1024];
  const int s = blockIdx.x;
  const T* a = A + (int64_t)s * ld;
  double acc = 0;
  for (int64_t n = threadIdx.x; n < N; n += blockDim.x) acc += W[n] * absq(widen(a[n]));
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[s] = 1.0 / sqrt(red[0]);
}

template <class T>
__global__ void colscale_kernel(T* __restrict__ A, const double* __restrict__ scal, int64_t N, int64_t ld, int ns) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (n >= N) return;
  T* a = A + (int64_t)s * ld;
  a[n] = narrow<T>(rmul(scal[s], widen(a[n])));
}

template <class D>
__global__ void hermitize_kernel(D* M, int ns) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (a >= ns || a > b) return;
  const D x = M[(int64_t)b * ns + a];
  const D y = M[(int64_t)a * ns + b];
  D m;
  if constexpr (traits<D>::is_complex) {
    m = D{0.5 * (x.x + y.x), 0.5 * (x.y - y.y)};
    M[(int64_t)b * ns + a] = m;
    M[(int64_t)a * ns + b] = D{m.x, -m.y};
  } else {
    m = 0.5 * (x + y);
    M[(int64_t)b * ns + a] = m;
    M[(int64_t)a * ns + b] = m;
  }
}

// ---------------------------------------------------------------- host side
template <class T>
static int gram_impl(cyx_ham_s* h, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                     const double* W, void* M_out) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  const int ta = ceil_div(na, GT), tb = ceil_div(nb, GT);
  int kparts = ceil_div(2048, ta * tb);
  int64_t krange = (N + kparts - 1) / kparts;
  krange = ((krange + GK - 1) / GK) * GK;
  kparts = (int)((N + krange - 1) / krange);
  const size_t pbytes = (size_t)kparts * na * nb * sizeof(D);
  CYX_CHECK(ensure(h->small, pbytes) == CYX_OK, "partials allocation failed");
  hipStream_t st = h->ctx->stream;
  PhaseTimer tm(h->ctx, PH_GRAM);
  hipLaunchKernelGGL(gram_partial_kernel<T>, dim3(ta, tb, kparts), dim3(256), 0, st, (const T*)A, lda, na,
                     (const T*)B, ldb, nb, W, N, krange, (D*)h->small.p);
  CYX_HIP(hipGetLastError());
  const int64_t len = (int64_t)na * nb;
  hipLaunchKernelGGL(reduce_partials_kernel<D>, dim3(ceil_div(len, 256)), dim3(256), 0, st, (const D*)h->small.p,
                     kparts, len, (D*)M_out);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int gram_rect(cyx_ham_s* h, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
              const double* W, void* M_out) {
  switch (dtype) {
    case CYX_F64:
      return gram_impl<double>(h, A, lda, na, B, ldb, nb, W, M_out);
    case CYX_F32:
      return gram_impl<float>(h, A, lda, na, B, ldb, nb, W, M_out);
    case CYX_C128:
      return gram_impl<cplx<double>>(h, A, lda, na, B, ldb, nb, W, M_out);
    case CYX_C64:
      return gram_impl<cplx<float>>(h, A, lda, na, B, ldb, nb, W, M_out);
  }
  return CYX_ERR_ARG;
}

int gram_w(cyx_ham_s* h, int dtype, const void* A, const void* B, int ns, int64_t ld, void* M_out) {
  return gram_rect(h, dtype, A, ld, ns, B, ld, ns, h->Wm, M_out);
}

template <class T>
static int rotate_impl(cyx_ham_s* h, const void* A, const void* Q, void* out, int ns, int64_t ld) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  PhaseTimer tm(h->ctx, PH_ROTATE);
  hipLaunchKernelGGL(rotate_kernel<T>, dim3(ceil_div(N, GT), ceil_div(ns, GT)), dim3(256), 0, h->ctx->stream,
                     (const T*)A, (const D*)Q, (T*)out, N, ld, ns);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int rotate(cyx_ham_s* h, int dtype, const void* A, const void* Q, void* out, int ns, int64_t ld) {
  switch (dtype) {
    case CYX_F64:
      return rotate_impl<double>(h, A, Q, out, ns, ld);
    case CYX_F32:
      return rotate_impl<float>(h, A, Q, out, ns, ld);
    case CYX_C128:
      return rotate_impl<cplx<double>>(h, A, Q, out, ns, ld);
    case CYX_C64:
      return rotate_impl<cplx<float>>(h, A, Q, out, ns, ld);
  }
  return CYX_ERR_ARG;
}

template <class T>
static int wnorm_impl(cyx_ham_s* h, void* A, int ns, int64_t ld, double* scal) {
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  PhaseTimer tm(h->ctx, PH_WNORM);
  hipLaunchKernelGGL(colnorm_kernel<T>, dim3(ns), dim3(1024), 0, st, (const T*)A, h->W, N, ld, scal);
  CYX_HIP(hipGetLastError());
  hipLaunchKernelGGL(colscale_kernel<T>, dim3(ceil_div(N, 256), ns), dim3(256), 0, st, (T*)A, (const double*)scal, N,
                     ld, ns);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int wnormalize(cyx_ham_s* h, int dtype, void* A, int ns, int64_t ld, double* scal) {
  switch (dtype) {
    case CYX_F64:
      return wnorm_impl<double>(h, A, ns, ld, scal);
    case CYX_F32:
      return wnorm_impl<float>(h, A, ns, ld, scal);
    case CYX_C128:
      return wnorm_impl<cplx<double>>(h, A, ns, ld, scal);
    case CYX_C64:
      return wnorm_impl<cplx<float>>(h, A, ns, ld, scal);
  }
  return CYX_ERR_ARG;
}

int hermitize(cyx_ham_s* h, int dtype, void* M, int ns) {
  hipStream_t st = h->ctx->stream;
  const bool cx = dtype == CYX_C128 || dtype == CYX_C64;
  if (cx)
    hipLaunchKernelGGL(hermitize_kernel<cplx<double>>, dim3(ceil_div(ns, 64), ns), dim3(64), 0, st, (cplx<double>*)M,
                       ns);
  else
    hipLaunchKernelGGL(hermitize_kernel<double>, dim3(ceil_div(ns, 64), ns), dim3(64), 0, st, (double*)M, ns);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

}  // namespace cyx
