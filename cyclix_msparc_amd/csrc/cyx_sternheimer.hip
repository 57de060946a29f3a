// Sternheimer operator and batched BiCGSTAB (src/phonon.m:2139-2157 and the
// solves at src/phonon.m:1516-1521).
//
//   A_j x = (H - lambda_j) x + Psi (Psi^H (W .* x))        one shift per column
//
// H is the channel's Hamiltonian (fused stencil + KB), Psi the nv occupied
// states, W the quadrature weights.  The batched solver runs BiCGSTAB
// (van der Vorst 1992, the iteration of MATLAB's bicgstab) on every column
// at once: each iteration makes two batched operator applications, and the
// per-column scalars (rho, alpha, omega, beta) live on the device, so the
// host only polls the number of active columns every few iterations.
// Columns that have converged keep their x and r (their alpha and omega are
// zero), so the batch can run on until the last column is done.
#include "cyx_internal.hpp"
#include "cyx_reduce.hpp"

namespace cyx {

__device__ __forceinline__ double cdiv(double a, double b) { return a / b; }
__device__ __forceinline__ cplx<double> cdiv(cplx<double> a, cplx<double> b) {
  const double d = b.x * b.x + b.y * b.y;
  return cplx<double>{(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__device__ __forceinline__ double re(double a) { return a; }
__device__ __forceinline__ double re(cplx<double> a) { return a.x; }
__device__ __forceinline__ double mag2(double a) { return a * a; }
__device__ __forceinline__ double mag2(cplx<double> a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double dscale(double a, double s) { return a * s; }
__device__ __forceinline__ cplx<double> dscale(cplx<double> a, double s) { return cplx<double>{a.x * s, a.y * s}; }
__device__ __forceinline__ double neg(double a) { return -a; }
__device__ __forceinline__ cplx<double> neg(cplx<double> a) { return cplx<double>{-a.x, -a.y}; }

template <class D>
struct ColState {
  D rho, rho_prev, alpha, omega, beta;
  double bnorm2, rnorm2;
  int active, half, iters, flag;  // flag: 0 ok, 1 not converged, 4 breakdown (rho or omega zero)
};

constexpr int SJB = 8;  // columns per thread in the finishing kernel

// HX[j] = HX[j] - lam[j] X[j] + sum_v PV[v] C[v, j]   (C column-major nv x ncol)
template <class T>
__global__ void __launch_bounds__(256) stern_finish_kernel(T* __restrict__ HX, const T* __restrict__ X, int64_t ld,
                                                           const T* __restrict__ PV, int64_t ldv, int nv,
                                                           const typename acc_of<T>::type* __restrict__ C,
                                                           const double* __restrict__ lam, int64_t N, int ncol) {
  using D = typename acc_of<T>::type;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int j0 = blockIdx.y * SJB;
  if (n >= N) return;
  D acc[SJB];
#pragma unroll
  for (int jj = 0; jj < SJB; ++jj) acc[jj] = D{};
  const int jn = min(SJB, ncol - j0);
  for (int v = 0; v < nv; ++v) {
    const D pv = widen(PV[(int64_t)v * ldv + n]);
#pragma unroll
    for (int jj = 0; jj < SJB; ++jj)
      if (jj < jn) acc[jj] = cfma(pv, C[(int64_t)(j0 + jj) * nv + v], acc[jj]);
  }
#pragma unroll
  for (int jj = 0; jj < SJB; ++jj) {
    if (jj < jn) {
      const int64_t o = (int64_t)(j0 + jj) * ld + n;
      const D hx = widen(HX[o]);
      const D x = widen(X[o]);
      HX[o] = narrow<T>(add(rfma(-lam[j0 + jj], x, hx), acc[jj]));
    }
  }
}

// ---------------------------------------------------------------- batched column reductions
// Partial sums over chunk c of column j of NP products conj(a_p) b_p, stored
// at part[(j * nchunk + c) * 2 + p].
template <class D>
__device__ __forceinline__ void store_partial(D v, double* red, D* part, int64_t slot) {
  if constexpr (traits<D>::is_complex) {
    const double a = block_sum(v.x, red);
    const double b = block_sum(v.y, red);
    if (threadIdx.x == 0) part[slot] = D{a, b};
  } else {
    const double a = block_sum(v, red);
    if (threadIdx.x == 0) part[slot] = a;
  }
}

struct ChunkRange {
  int64_t n0, n1;
};
__device__ __forceinline__ ChunkRange chunk_of(int64_t N, int nchunk) {
  const int64_t per = (N + nchunk - 1) / nchunk;
  const int64_t n0 = (int64_t)blockIdx.x * per;
  return ChunkRange{n0, min(N, n0 + per)};
}

// part <- (a0, b0) [, (a1, b1)] column dots
template <class T, int NP>
__global__ void __launch_bounds__(AB) coldot_kernel(const T* __restrict__ a0, const T* __restrict__ b0,
                                                    const T* __restrict__ a1, const T* __restrict__ b1, int64_t N,
                                                    int64_t ld, int nchunk, typename acc_of<T>::type* part) {
  using D = typename acc_of<T>::type;
  __shared__ double red[AB];
  const int j = blockIdx.y;
  const ChunkRange cr = chunk_of(N, nchunk);
  D s0{}, s1{};
  for (int64_t n = cr.n0 + threadIdx.x; n < cr.n1; n += AB) {
    const int64_t o = (int64_t)j * ld + n;
    s0 = cjfma(widen(a0[o]), widen(b0[o]), s0);
    if constexpr (NP == 2) s1 = cjfma(widen(a1[o]), widen(b1[o]), s1);
  }
  const int64_t slot = ((int64_t)j * nchunk + blockIdx.x) * 2;
  store_partial(s0, red, part, slot);
  if constexpr (NP == 2) store_partial(s1, red, part, slot + 1);
}

// fixed-order sum of the nchunk partials of column blockIdx.x, pair p
template <class D>
__device__ D sum_partials(const D* part, int nchunk, int p, double* red) {
  const int j = blockIdx.x;
  D s{};
  for (int c = threadIdx.x; c < nchunk; c += AB) s = add(s, part[((int64_t)j * nchunk + c) * 2 + p]);
  if constexpr (traits<D>::is_complex) {
    const double a = block_sum(s.x, red);
    const double b = block_sum(s.y, red);
    return D{a, b};
  } else {
    return block_sum(s, red);
  }
}

// ---------------------------------------------------------------- BiCGSTAB stage kernels
// r = b - v (v = A x), rt = r, p = v = 0; partials ||b||^2, (rt, r)
template <class T>
__global__ void __launch_bounds__(AB) bi_init_kernel(const T* __restrict__ b, int64_t ldb, T* __restrict__ r,
                                                     T* __restrict__ rt, T* __restrict__ p, T* __restrict__ v,
                                                     int64_t N, int nchunk, typename acc_of<T>::type* part) {
  using D = typename acc_of<T>::type;
  __shared__ double red[AB];
  const int j = blockIdx.y;
  const ChunkRange cr = chunk_of(N, nchunk);
  D sb{}, sr{};
  for (int64_t n = cr.n0 + threadIdx.x; n < cr.n1; n += AB) {
    const int64_t o = (int64_t)j * N + n;
    const D bv = widen(b[(int64_t)j * ldb + n]);
    const D rv = sub(bv, widen(v[o]));
    const T rT = narrow<T>(rv);
    r[o] = rT;
    rt[o] = rT;
    p[o] = T{};
    v[o] = T{};
    sb = cjfma(bv, bv, sb);
    sr = cjfma(rv, rv, sr);
  }
  const int64_t slot = ((int64_t)j * nchunk + blockIdx.x) * 2;
  store_partial(sb, red, part, slot);
  store_partial(sr, red, part, slot + 1);
}

template <class D>
__global__ void __launch_bounds__(AB) bi_init_finalize(const D* part, int nchunk, double tol, ColState<D>* S) {
  __shared__ double red[AB];
  const D bb = sum_partials(part, nchunk, 0, red);
  const D rr = sum_partials(part, nchunk, 1, red);
  if (threadIdx.x == 0) {
    ColState<D>& c = S[blockIdx.x];
    c.bnorm2 = re(bb);
    c.rnorm2 = re(rr);
    c.rho = rr;
    c.rho_prev = rr;
    c.alpha = D{};
    c.omega = D{};
    // the first p = r + beta (p - omega v) with p = v = 0 is r for any beta
    c.beta = D{};
    c.iters = 0;
    c.half = 0;
    c.flag = 0;
    c.active = (c.bnorm2 > 0.0 && sqrt(c.rnorm2) > tol * sqrt(c.bnorm2)) ? 1 : 0;
  }
}

// p = r + beta (p - omega v);  ph = M^{-1} p   (M: per-radius diagonal)
template <class T>
__global__ void __launch_bounds__(256) bi_p_kernel(const T* __restrict__ r, T* __restrict__ p,
                                                   const T* __restrict__ v, T* __restrict__ ph,
                                                   const double* __restrict__ dinv, int nr, int64_t N,
                                                   const ColState<typename acc_of<T>::type>* S) {
  using D = typename acc_of<T>::type;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (n >= N || !S[j].active) return;
  const D beta = S[j].beta, omega = S[j].omega;
  const int64_t o = (int64_t)j * N + n;
  const D pn = add(widen(r[o]), cmul(beta, sub(widen(p[o]), cmul(omega, widen(v[o])))));
  p[o] = narrow<T>(pn);
  ph[o] = narrow<T>(dinv ? rmul(dinv[n % nr], pn) : pn);
}

// alpha = rho / (rt, v)
template <class D>
__global__ void __launch_bounds__(AB) bi_alpha_finalize(const D* part, int nchunk, ColState<D>* S) {
  __shared__ double red[AB];
  const D rtv = sum_partials(part, nchunk, 0, red);
  if (threadIdx.x == 0) {
    ColState<D>& c = S[blockIdx.x];
    if (c.active) {
      if (mag2(rtv) == 0.0) {
        c.alpha = D{};
        c.active = 0;
        c.flag = 4;
      } else {
        c.alpha = cdiv(c.rho, rtv);
      }
    }
  }
}

// s = r - alpha v (into r);  sh = M^{-1} s;  partial ||s||^2
template <class T>
__global__ void __launch_bounds__(AB) bi_s_kernel(T* __restrict__ r, const T* __restrict__ v, T* __restrict__ sh,
                                                  const double* __restrict__ dinv, int nr, int64_t N, int nchunk,
                                                  const ColState<typename acc_of<T>::type>* S,
                                                  typename acc_of<T>::type* part) {
  using D = typename acc_of<T>::type;
  __shared__ double red[AB];
  const int j = blockIdx.y;
  const ChunkRange cr = chunk_of(N, nchunk);
  const bool act = S[j].active;
  const D alpha = S[j].alpha;
  double ss = 0.0;
  for (int64_t n = cr.n0 + threadIdx.x; n < cr.n1; n += AB) {
    const int64_t o = (int64_t)j * N + n;
    D sv = widen(r[o]);
    if (act) {
      sv = sub(sv, cmul(alpha, widen(v[o])));
      r[o] = narrow<T>(sv);
      sh[o] = narrow<T>(dinv ? rmul(dinv[n % nr], sv) : sv);
    }
    ss += mag2(sv);
  }
  const double tot = block_sum(ss, red);
  if (threadIdx.x == 0) part[((int64_t)j * nchunk + blockIdx.x) * 2] = D{tot};
}

// half-step test (MATLAB's bicgstab stops at x + alpha ph when ||s|| <= tol ||b||)
template <class D>
__global__ void __launch_bounds__(AB) bi_s_finalize(const D* part, int nchunk, double tol, ColState<D>* S) {
  __shared__ double red[AB];
  const D ss = sum_partials(part, nchunk, 0, red);
  if (threadIdx.x == 0) {
    ColState<D>& c = S[blockIdx.x];
    if (c.active && sqrt(re(ss)) <= tol * sqrt(c.bnorm2)) c.half = 1;
  }
}

// omega = (t, s) / (t, t)   (0 for a half-step exit)
template <class D>
__global__ void __launch_bounds__(AB) bi_omega_finalize(const D* part, int nchunk, ColState<D>* S) {
  __shared__ double red[AB];
  const D ts = sum_partials(part, nchunk, 0, red);
  const D tt = sum_partials(part, nchunk, 1, red);
  if (threadIdx.x == 0) {
    ColState<D>& c = S[blockIdx.x];
    if (c.active) {
      if (c.half) {
        c.omega = D{};
      } else if (re(tt) == 0.0) {
        c.omega = D{};
        c.flag = 4;
      } else {
        c.omega = cdiv(ts, tt);
      }
    }
  }
}

// x += alpha ph + omega sh;  r = s - omega t;  partials (rt, r), ||r||^2
template <class T>
__global__ void __launch_bounds__(AB) bi_x_kernel(T* __restrict__ x, int64_t ldx, T* __restrict__ r,
                                                  const T* __restrict__ rt, const T* __restrict__ ph,
                                                  const T* __restrict__ sh, const T* __restrict__ t, int64_t N,
                                                  int nchunk, const ColState<typename acc_of<T>::type>* S,
                                                  typename acc_of<T>::type* part) {
  using D = typename acc_of<T>::type;
  __shared__ double red[AB];
  const int j = blockIdx.y;
  const ChunkRange cr = chunk_of(N, nchunk);
  const bool act = S[j].active;
  const D alpha = S[j].alpha, omega = S[j].omega;
  D srr{};
  double s2 = 0.0;
  for (int64_t n = cr.n0 + threadIdx.x; n < cr.n1; n += AB) {
    const int64_t o = (int64_t)j * N + n;
    D rv = widen(r[o]);
    if (act) {
      const int64_t ox = (int64_t)j * ldx + n;
      D xv = widen(x[ox]);
      xv = add(xv, add(cmul(alpha, widen(ph[o])), cmul(omega, widen(sh[o]))));
      x[ox] = narrow<T>(xv);
      rv = sub(rv, cmul(omega, widen(t[o])));
      r[o] = narrow<T>(rv);
    }
    srr = cjfma(widen(rt[o]), rv, srr);
    s2 += mag2(rv);
  }
  const int64_t slot = ((int64_t)j * nchunk + blockIdx.x) * 2;
  store_partial(srr, red, part, slot);
  const double tot = block_sum(s2, red);
  if (threadIdx.x == 0) part[slot + 1] = D{tot};
}

template <class D>
__global__ void __launch_bounds__(AB) bi_r_finalize(const D* part, int nchunk, double tol, ColState<D>* S) {
  __shared__ double red[AB];
  const D rr = sum_partials(part, nchunk, 0, red);
  const D r2 = sum_partials(part, nchunk, 1, red);
  if (threadIdx.x == 0) {
    ColState<D>& c = S[blockIdx.x];
    if (c.active) {
      c.iters += 1;
      c.rnorm2 = re(r2);
      c.rho_prev = c.rho;
      c.rho = rr;
      if (c.half || sqrt(c.rnorm2) <= tol * sqrt(c.bnorm2)) {
        c.active = 0;
      } else if (c.flag == 4 || mag2(c.rho) == 0.0 || mag2(c.omega) == 0.0) {
        c.active = 0;
        c.flag = 4;
      } else {
        c.beta = cmul(cdiv(c.rho, c.rho_prev), cdiv(c.alpha, c.omega));
      }
    }
  }
}

template <class D>
__global__ void count_active_kernel(const ColState<D>* S, int ncol, int* out) {
  __shared__ double red[AB];
  double a = 0;
  for (int j = threadIdx.x; j < ncol; j += AB) a += S[j].active;
  const double tot = block_sum(a, red);
  if (threadIdx.x == 0) *out = (int)tot;
}

// per-radius inverse diagonal of -1/2 Lap (the Jacobi preconditioner)
__global__ void dinv_kernel(const double* rcoef, int nr, double* dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nr) dinv[i] = 1.0 / rcoef[(int64_t)i * RCOEF + RCOEF - 1];
}

// ---------------------------------------------------------------- host side
struct SternLayout {
  size_t blk;  // bytes of one ncol x N block
  char* base;
  void* r;
  void* rt;
  void* p;
  void* v;
  void* ph;
  void* sh;
  void* t;
  void* C;
  double* lam;
  void* part;
  void* state;
  double* dinv;
  int* count;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <class T>
static int stern_layout(cyx_ham_s* h, int ncol, int nv, int nchunk, bool solver, SternLayout& L) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  L.blk = align256((size_t)ncol * N * sizeof(T));
  const size_t nblk = solver ? 7 : 0;
  const size_t bC = align256((size_t)nv * ncol * sizeof(D) + 16);
  const size_t bl = align256((size_t)ncol * sizeof(double));
  const size_t bp = align256((size_t)ncol * nchunk * 2 * sizeof(D));
  const size_t bs = align256((size_t)ncol * sizeof(ColState<D>));
  const size_t bd = align256((size_t)h->g.nr * sizeof(double));
  const size_t total = nblk * L.blk + bC + bl + bp + bs + bd + 256;
  if (ensure(h->stern, total) != CYX_OK) return CYX_ERR_NOMEM;
  char* b = (char*)h->stern.p;
  L.base = b;
  void** blocks[7] = {&L.r, &L.rt, &L.p, &L.v, &L.ph, &L.sh, &L.t};
  for (size_t k = 0; k < 7; ++k) *blocks[k] = k < nblk ? (void*)(b + k * L.blk) : nullptr;
  b += nblk * L.blk;
  L.C = b;
  b += bC;
  L.lam = (double*)b;
  b += bl;
  L.part = b;
  b += bp;
  L.state = b;
  b += bs;
  L.dinv = (double*)b;
  b += bd;
  L.count = (int*)b;
  return CYX_OK;
}

// Y = A X for all columns (Y != X)
template <class T>
static int stern_op(cyx_ham_s* h, int dtype, const T* X, int64_t ldx, T* Y, int64_t ldy, int ncol, const T* PV,
                    int64_t ldv, int nv, const SternLayout& L) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  int rc;
  // the finishing kernel reads X at Y's leading dimension
  if (ldx != ldy) {
    set_error("internal: Sternheimer operator needs equal leading dimensions");
    return CYX_ERR_ARG;
  }
  if ((rc = stencil_step(h, dtype, X, nullptr, Y, ncol, ldy, 1.0, 0.0, 0.0)) != CYX_OK) return rc;
  if (nv > 0) {
    if ((rc = gram_rect(h, dtype, PV, ldv, nv, X, ldx, ncol, h->W, L.C)) != CYX_OK) return rc;
  }
  hipLaunchKernelGGL(stern_finish_kernel<T>, dim3(ceil_div(N, 256), ceil_div(ncol, SJB)), dim3(256), 0,
                     h->ctx->stream, Y, X, ldy, PV, ldv, nv, (const D*)L.C, (const double*)L.lam, N, ncol);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

template <class T>
static int stern_apply_impl(cyx_ham_s* h, int dtype, const void* X, void* Y, int ncol, int64_t ld,
                            const double* lam, const void* PV, int64_t ldv, int nv) {
  SternLayout L;
  CYX_CHECK(stern_layout<T>(h, ncol, nv, 1, false, L) == CYX_OK, "Sternheimer workspace allocation failed");
  CYX_HIP(hipMemcpyAsync(L.lam, lam, ncol * sizeof(double), hipMemcpyHostToDevice, h->ctx->stream));
  return stern_op<T>(h, dtype, (const T*)X, ld, (T*)Y, ld, ncol, (const T*)PV, ldv, nv, L);
}

template <class T>
static int stern_solve_impl(cyx_ham_s* h, int dtype, const void* B, void* X, int ncol, int64_t ld, const double* lam,
                            const void* PV, int64_t ldv, int nv, int use_guess, double tol, int maxit,
                            int precond, cyx_stern_info* info, int32_t* col_iters, double* col_relres) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(128, ceil_div(N, 2048)));
  SternLayout L;
  CYX_CHECK(stern_layout<T>(h, ncol, nv, nchunk, true, L) == CYX_OK, "Sternheimer workspace allocation failed");
  CYX_HIP(hipMemcpyAsync(L.lam, lam, ncol * sizeof(double), hipMemcpyHostToDevice, st));
  const double* dinv = nullptr;
  if (precond) {
    hipLaunchKernelGGL(dinv_kernel, dim3(ceil_div(h->g.nr, 256)), dim3(256), 0, st, (const double*)h->c64.rcoef,
                       h->g.nr, L.dinv);
    CYX_HIP(hipGetLastError());
    dinv = L.dinv;
  }
  T* x = (T*)X;
  T *r = (T*)L.r, *rt = (T*)L.rt, *p = (T*)L.p, *v = (T*)L.v, *ph = (T*)L.ph, *sh = (T*)L.sh, *t = (T*)L.t;
  D* part = (D*)L.part;
  ColState<D>* S = (ColState<D>*)L.state;
  const T* pv = (const T*)PV;
  int rc;
  // x0 (zero unless a guess is given); v = A x0 (bicgstab's r = b - A x0)
  if (!use_guess) CYX_HIP(hipMemset2DAsync(x, ld * sizeof(T), 0, N * sizeof(T), ncol, st));
  CYX_HIP(hipMemcpy2DAsync(ph, N * sizeof(T), x, ld * sizeof(T), N * sizeof(T), ncol, hipMemcpyDeviceToDevice, st));
  if ((rc = stern_op<T>(h, dtype, ph, N, v, N, ncol, pv, ldv, nv, L)) != CYX_OK) return rc;
  const dim3 gc(nchunk, ncol), gp(ceil_div(N, 256), ncol);
  hipLaunchKernelGGL(bi_init_kernel<T>, gc, dim3(AB), 0, st, (const T*)B, ld, r, rt, p, v, N, nchunk, part);
  hipLaunchKernelGGL(bi_init_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, tol, S);
  CYX_HIP(hipGetLastError());
  int active = ncol, it = 0;
  const int poll = 4;
  while (it < maxit && active > 0) {
    hipLaunchKernelGGL(bi_p_kernel<T>, gp, dim3(256), 0, st, (const T*)r, p, (const T*)v, ph, dinv, h->g.nr, N,
                       (const ColState<D>*)S);
    if ((rc = stern_op<T>(h, dtype, ph, N, v, N, ncol, pv, ldv, nv, L)) != CYX_OK) return rc;
    hipLaunchKernelGGL((coldot_kernel<T, 1>), gc, dim3(AB), 0, st, (const T*)rt, (const T*)v, (const T*)nullptr,
                       (const T*)nullptr, N, N, nchunk, part);
    hipLaunchKernelGGL(bi_alpha_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, S);
    hipLaunchKernelGGL(bi_s_kernel<T>, gc, dim3(AB), 0, st, r, (const T*)v, sh, dinv, h->g.nr, N, nchunk,
                       (const ColState<D>*)S, part);
    hipLaunchKernelGGL(bi_s_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, tol, S);
    if ((rc = stern_op<T>(h, dtype, sh, N, t, N, ncol, pv, ldv, nv, L)) != CYX_OK) return rc;
    hipLaunchKernelGGL((coldot_kernel<T, 2>), gc, dim3(AB), 0, st, (const T*)t, (const T*)r, (const T*)t,
                       (const T*)t, N, N, nchunk, part);
    hipLaunchKernelGGL(bi_omega_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, S);
    hipLaunchKernelGGL(bi_x_kernel<T>, gc, dim3(AB), 0, st, x, ld, r, (const T*)rt, (const T*)ph, (const T*)sh,
                       (const T*)t, N, nchunk, (const ColState<D>*)S, part);
    hipLaunchKernelGGL(bi_r_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, tol, S);
    CYX_HIP(hipGetLastError());
    ++it;
    if (it % poll == 0 || it == maxit) {
      hipLaunchKernelGGL(count_active_kernel<D>, dim3(1), dim3(AB), 0, st, (const ColState<D>*)S, ncol, L.count);
      CYX_HIP(hipMemcpyAsync(&active, L.count, sizeof(int), hipMemcpyDeviceToHost, st));
      CYX_HIP(hipStreamSynchronize(st));
    }
  }
  std::vector<ColState<D>> hs(ncol);
  CYX_HIP(hipMemcpyAsync(hs.data(), S, ncol * sizeof(ColState<D>), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  int nconv = 0, maxits = 0, breakdown = 0;
  double maxrel = 0.0;
  for (int j = 0; j < ncol; ++j) {
    const double rel = hs[j].bnorm2 > 0 ? std::sqrt(hs[j].rnorm2 / hs[j].bnorm2) : 0.0;
    // a half-step exit leaves r = s, whose norm met the tolerance
    const bool ok = hs[j].bnorm2 == 0.0 || hs[j].half || rel <= tol;
    nconv += ok;
    breakdown += hs[j].flag == 4 && !ok;
    maxits = std::max(maxits, hs[j].iters);
    maxrel = std::max(maxrel, rel);
    if (col_iters) col_iters[j] = hs[j].iters;
    if (col_relres) col_relres[j] = rel;
  }
  if (info) {
    info->max_iterations = maxits;
    info->n_converged = nconv;
    info->max_relres = maxrel;
  }
  if (breakdown) {
    set_error("BiCGSTAB broke down (rho or omega zero) on %d of %d columns", breakdown, ncol);
    return CYX_ERR_NOCONV;
  }
  if (nconv < ncol) {
    set_error("BiCGSTAB reached %d iterations with %d of %d columns above tol %g", maxit, ncol - nconv, ncol, tol);
    return CYX_ERR_NOCONV;
  }
  return CYX_OK;
}

}  // namespace cyx

using namespace cyx;

static int stern_check(cyx_ham_t h, int dtype, int ncol, int64_t ld, const double* lam, const void* PV, int64_t ldv,
                       int nv) {
  CYX_CHECK(h && lam, "null argument");
  CYX_CHECK(dtype == CYX_F64 || dtype == CYX_C128 || dtype == CYX_F32 || dtype == CYX_C64, "bad dtype %d", dtype);
  CYX_CHECK(!h->bloch || dtype == CYX_C128 || dtype == CYX_C64, "a Bloch channel needs a complex dtype");
  CYX_CHECK(ncol >= 1 && ld >= h->N(), "bad block shape (ncol=%d ld=%lld N=%lld)", ncol, (long long)ld,
            (long long)h->N());
  CYX_CHECK(nv >= 0 && (nv == 0 || (PV && ldv >= h->N())), "bad projector block (nv=%d)", nv);
  return CYX_OK;
}

extern "C" {

int cyx_sternheimer_apply(cyx_ham_t h, int dtype, const void* X, void* Y, int ncol, int64_t ld, const double* lam,
                          const void* psi_v, int64_t ldv, int nv) {
  int rc;
  if ((rc = stern_check(h, dtype, ncol, ld, lam, psi_v, ldv, nv)) != CYX_OK) return rc;
  CYX_CHECK(X && Y && X != Y, "null or aliased argument");
  switch (dtype) {
    case CYX_F64: return stern_apply_impl<double>(h, dtype, X, Y, ncol, ld, lam, psi_v, ldv, nv);
    case CYX_F32: return stern_apply_impl<float>(h, dtype, X, Y, ncol, ld, lam, psi_v, ldv, nv);
    case CYX_C128: return stern_apply_impl<cplx<double>>(h, dtype, X, Y, ncol, ld, lam, psi_v, ldv, nv);
    default: return stern_apply_impl<cplx<float>>(h, dtype, X, Y, ncol, ld, lam, psi_v, ldv, nv);
  }
}

int cyx_sternheimer_solve(cyx_ham_t h, int dtype, const void* B, void* X, int ncol, int64_t ld, const double* lam,
                          const void* psi_v, int64_t ldv, int nv, int use_guess, double tol, int max_iter,
                          int precond, cyx_stern_info* info, int32_t* col_iters, double* col_relres) {
  int rc;
  if ((rc = stern_check(h, dtype, ncol, ld, lam, psi_v, ldv, nv)) != CYX_OK) return rc;
  CYX_CHECK(B && X && B != X, "null or aliased argument");
  CYX_CHECK(tol > 0 && max_iter >= 1, "bad tolerance or iteration limit");
  switch (dtype) {
    case CYX_F64:
      return stern_solve_impl<double>(h, dtype, B, X, ncol, ld, lam, psi_v, ldv, nv, use_guess, tol, max_iter,
                                      precond, info, col_iters, col_relres);
    case CYX_F32:
      return stern_solve_impl<float>(h, dtype, B, X, ncol, ld, lam, psi_v, ldv, nv, use_guess, tol, max_iter,
                                     precond, info, col_iters, col_relres);
    case CYX_C128:
      return stern_solve_impl<cplx<double>>(h, dtype, B, X, ncol, ld, lam, psi_v, ldv, nv, use_guess, tol,
                                            max_iter, precond, info, col_iters, col_relres);
    default:
      return stern_solve_impl<cplx<float>>(h, dtype, B, X, ncol, ld, lam, psi_v, ldv, nv, use_guess, tol, max_iter,
                                           precond, info, col_iters, col_relres);
  }
}

}  // extern "C"
