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

constexpr int SJB = 16;  // columns per thread in the finishing kernel (psi_v is read once per SJB columns)

// HX[j] = HX[j] - lam[j] X[j] + sum_v PV[v] C[v, j]   (C column-major nv x ncol);
// SHIFT = false: HX[j] += sum_v PV[v] C[v, j] (X and lam unused)
template <class T, bool SHIFT = true>
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
      if constexpr (SHIFT) {
        const D x = widen(X[o]);
        HX[o] = narrow<T>(add(rfma(-lam[j0 + jj], x, hx), acc[jj]));
      } else {
        HX[o] = narrow<T>(add(hx, acc[jj]));
      }
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

// F = coeff .* (-G) and dlam_n = -Re G_nn, where G = psi^H W B with B = -dH psi
// (so -G is psi_DH_psi of src/phonon.m:1360,1402-1405)
template <class D>
__global__ void __launch_bounds__(256) rhs_coef_kernel(const D* __restrict__ G, const double* __restrict__ coeff,
                                                       D* __restrict__ F, double* __restrict__ dlam, int nv) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= nv * nv) return;
  const D dh = neg(G[idx]);
  F[idx] = dscale(dh, coeff[idx]);
  if (idx % nv == idx / nv) dlam[idx / nv] = re(dh);
}

// ---------------------------------------------------------------- host side
struct SternLayout {
  size_t blk = 0;  // bytes of one ncol x N block
  void* r = nullptr;
  void* rt = nullptr;
  void* p = nullptr;
  void* v = nullptr;
  void* ph = nullptr;
  void* sh = nullptr;
  void* t = nullptr;
  void* C = nullptr;
  double* lam = nullptr;
  void* part = nullptr;
  void* state = nullptr;
  double* dinv = nullptr;
  int* count = nullptr;
  // KB coefficients / partials of the stencil steps and the Gram partials:
  // in the workspace when `own` (a captured graph must not use the handle's
  // lazily grown buffers), else null (the handle's)
  bool own = false;
  KbScratch ks{nullptr, nullptr};
  void* gpart = nullptr;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace of ncol columns against nv occupied states: the seven BiCGSTAB
// blocks (solver only), the projections C, the shifts, the column-dot
// partials, the column states, the Jacobi diagonal, the active count and,
// when own, the step scratch.  Returns the bytes; assigns the pointers when
// base is not null.
template <class T>
static size_t stern_layout(cyx_ham_s* h, int dtype, int ncol, int nv, int nchunk, bool solver, bool own, char* base,
                           SternLayout& L) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  size_t off = 0;
  const auto take = [&](size_t bytes) -> char* {
    char* q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  L.blk = align256((size_t)ncol * N * sizeof(T));
  void** blocks[7] = {&L.r, &L.rt, &L.p, &L.v, &L.ph, &L.sh, &L.t};
  for (void** b : blocks) *b = solver ? take(L.blk) : nullptr;
  L.C = take((size_t)nv * ncol * sizeof(D) + 16);
  L.lam = (double*)take((size_t)ncol * sizeof(double));
  L.part = take((size_t)ncol * nchunk * 2 * sizeof(D));
  L.state = take((size_t)ncol * sizeof(ColState<D>));
  L.dinv = (double*)take((size_t)h->g.nr * sizeof(double));
  L.count = (int*)take(sizeof(int));
  L.own = own;
  L.ks = KbScratch{nullptr, nullptr};
  L.gpart = nullptr;
  if (own) {
    // kb_scratch_bytes' layout: C rounded up to 256 bytes, then the partials
    const size_t kb = kb_scratch_bytes(h, dtype, ncol);
    char* k = take(std::max<size_t>(kb, 1));
    if (k) L.ks = KbScratch{k, k + align256((size_t)ncol * h->nproj * sizeof(T))};
    if (nv > 0) L.gpart = take(gram_part_bytes(dtype, N, nv, ncol, false));
  }
  return off + 256;
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
  if ((rc = stencil_step(h, dtype, X, nullptr, Y, ncol, ldy, 1.0, 0.0, 0.0, L.own ? &L.ks : nullptr)) != CYX_OK)
    return rc;
  if (nv > 0) {
    if ((rc = gram_rect(h, dtype, PV, ldv, nv, X, ldx, ncol, h->W, L.C, false, L.gpart)) != CYX_OK) return rc;
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
  const size_t bytes = stern_layout<T>(h, dtype, ncol, nv, 1, false, false, nullptr, L);
  CYX_CHECK(ensure(h->stern, bytes) == CYX_OK, "Sternheimer workspace allocation failed");
  stern_layout<T>(h, dtype, ncol, nv, 1, false, false, (char*)h->stern.p, L);
  CYX_HIP(hipMemcpyAsync(L.lam, lam, ncol * sizeof(double), hipMemcpyHostToDevice, h->ctx->stream));
  return stern_op<T>(h, dtype, (const T*)X, ld, (T*)Y, ld, ncol, (const T*)PV, ldv, nv, L);
}

// B = coeff-weighted occupied part - dH psi (src/phonon.m:1440-1444), see cyx_sternheimer_rhs
template <class T>
static int stern_rhs_impl(cyx_ham_s* h, int dtype, const void* psi, int64_t ldpsi, int nv, const double* coeff,
                          void* B, int64_t ldb, double* dlambda) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  const size_t bG = align256((size_t)nv * nv * sizeof(D));
  const size_t bc = align256((size_t)nv * nv * sizeof(double));
  const size_t bl = align256((size_t)nv * sizeof(double));
  CYX_CHECK(ensure(h->stern, 2 * bG + bc + bl) == CYX_OK, "Sternheimer workspace allocation failed");
  char* b = (char*)h->stern.p;
  D* G = (D*)b;
  D* F = (D*)(b + bG);
  double* cf = (double*)(b + 2 * bG);
  double* dl = (double*)(b + 2 * bG + bc);
  int rc;
  // B = -dH psi
  if ((rc = potential_apply(h, dtype, psi, ldpsi, B, ldb, nv, -1.0)) != CYX_OK) return rc;
  // G = psi^H W B = -psi_DH_psi
  if ((rc = gram_rect(h, dtype, psi, ldpsi, nv, B, ldb, nv, h->W, G, false)) != CYX_OK) return rc;
  CYX_HIP(hipMemcpyAsync(cf, coeff, (size_t)nv * nv * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(rhs_coef_kernel<D>, dim3(ceil_div(nv * nv, 256)), dim3(256), 0, st, (const D*)G,
                     (const double*)cf, F, dl, nv);
  // B_n += sum_m psi_m F_mn
  hipLaunchKernelGGL((stern_finish_kernel<T, false>), dim3(ceil_div(N, 256), ceil_div(nv, SJB)), dim3(256), 0, st,
                     (T*)B, (const T*)nullptr, ldb, (const T*)psi, ldpsi, nv, (const D*)F, (const double*)nullptr, N,
                     nv);
  CYX_HIP(hipGetLastError());
  if (dlambda) {
    CYX_HIP(hipMemcpyAsync(dlambda, dl, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, st));
    CYX_HIP(hipStreamSynchronize(st));
  }
  return CYX_OK;
}

}  // namespace cyx

// A batched BiCGSTAB over ncol columns: the workspace, the caller's blocks,
// and the captured graph of `graph_iters` iterations.
struct cyx_stern_plan_s {
  cyx_ham_s* h = nullptr;
  int dtype = 0, ncol = 0, nv = 0, nchunk = 1;
  const void* B = nullptr;
  void* X = nullptr;
  int64_t ld = 0;
  const void* PV = nullptr;
  int64_t ldv = 0;
  double tol = 0.0;
  bool precond = false;   // Jacobi diagonal of -1/2 Lap (precond = 1)
  bool spectral = false;  // exact Lap_std^-1 (precond = 2): eager iterations, no graph
  bool owns_ws = false;  // workspace from hipMalloc (public plans) or the handle's buffer (one solve)
  void* ws = nullptr;
  cyx::SternLayout L;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_iters = 0;
  hipStream_t cap = nullptr;
};

namespace cyx {

template <class T>
static int stern_init(cyx_stern_plan_s* sp, const double* lam, int use_guess) {
  using D = typename acc_of<T>::type;
  cyx_ham_s* h = sp->h;
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  const SternLayout& L = sp->L;
  const int ncol = sp->ncol, nchunk = sp->nchunk;
  CYX_HIP(hipMemcpyAsync(L.lam, lam, ncol * sizeof(double), hipMemcpyHostToDevice, st));
  if (sp->precond) {
    hipLaunchKernelGGL(dinv_kernel, dim3(ceil_div(h->g.nr, 256)), dim3(256), 0, st, (const double*)h->c64.rcoef,
                       h->g.nr, L.dinv);
    CYX_HIP(hipGetLastError());
  }
  T* x = (T*)sp->X;
  // x0 (zero unless a guess is given); v = A x0 (bicgstab's r = b - A x0)
  if (!use_guess) CYX_HIP(hipMemset2DAsync(x, sp->ld * sizeof(T), 0, N * sizeof(T), ncol, st));
  CYX_HIP(hipMemcpy2DAsync(L.ph, N * sizeof(T), x, sp->ld * sizeof(T), N * sizeof(T), ncol, hipMemcpyDeviceToDevice,
                           st));
  int rc;
  if ((rc = stern_op<T>(h, sp->dtype, (const T*)L.ph, N, (T*)L.v, N, ncol, (const T*)sp->PV, sp->ldv, sp->nv, L)) !=
      CYX_OK)
    return rc;
  const dim3 gc(nchunk, ncol);
  hipLaunchKernelGGL(bi_init_kernel<T>, gc, dim3(AB), 0, st, (const T*)sp->B, sp->ld, (T*)L.r, (T*)L.rt, (T*)L.p,
                     (T*)L.v, N, nchunk, (D*)L.part);
  hipLaunchKernelGGL(bi_init_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)L.part, nchunk, sp->tol,
                     (ColState<D>*)L.state);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

// one BiCGSTAB iteration of every active column on the context's stream
template <class T>
static int stern_iteration(cyx_stern_plan_s* sp) {
  using D = typename acc_of<T>::type;
  cyx_ham_s* h = sp->h;
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  const SternLayout& L = sp->L;
  const int ncol = sp->ncol, nchunk = sp->nchunk, dtype = sp->dtype;
  const double* dinv = sp->precond ? L.dinv : nullptr;
  T* x = (T*)sp->X;
  // M^-1 z = Lap_std^-1 z (the reference's ILU factors of Lap_std, src/phonon.m:1519)
  auto spectral = [&](const T* z, T* out) -> int {
    if constexpr (std::is_same_v<T, double>)
      return spectral_solve_real(h->spec, z, out, ncol, N, st);
    else if constexpr (std::is_same_v<T, cplx<double>>)
      return spectral_solve_complex(h->spec, z, out, ncol, N, st);
    else
      return CYX_ERR_ARG;
  };
  T *r = (T*)L.r, *rt = (T*)L.rt, *p = (T*)L.p, *v = (T*)L.v, *ph = (T*)L.ph, *sh = (T*)L.sh, *t = (T*)L.t;
  D* part = (D*)L.part;
  ColState<D>* S = (ColState<D>*)L.state;
  const T* pv = (const T*)sp->PV;
  const dim3 gc(nchunk, ncol), gp(ceil_div(N, 256), ncol);
  int rc;
  hipLaunchKernelGGL(bi_p_kernel<T>, gp, dim3(256), 0, st, (const T*)r, p, (const T*)v, ph, dinv, h->g.nr, N,
                     (const ColState<D>*)S);
  if (sp->spectral && (rc = spectral(p, ph)) != CYX_OK) return rc;
  if ((rc = stern_op<T>(h, dtype, ph, N, v, N, ncol, pv, sp->ldv, sp->nv, L)) != CYX_OK) return rc;
  hipLaunchKernelGGL((coldot_kernel<T, 1>), gc, dim3(AB), 0, st, (const T*)rt, (const T*)v, (const T*)nullptr,
                     (const T*)nullptr, N, N, nchunk, part);
  hipLaunchKernelGGL(bi_alpha_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, S);
  hipLaunchKernelGGL(bi_s_kernel<T>, gc, dim3(AB), 0, st, r, (const T*)v, sh, dinv, h->g.nr, N, nchunk,
                     (const ColState<D>*)S, part);
  hipLaunchKernelGGL(bi_s_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, sp->tol, S);
  if (sp->spectral && (rc = spectral(r, sh)) != CYX_OK) return rc;
  if ((rc = stern_op<T>(h, dtype, sh, N, t, N, ncol, pv, sp->ldv, sp->nv, L)) != CYX_OK) return rc;
  hipLaunchKernelGGL((coldot_kernel<T, 2>), gc, dim3(AB), 0, st, (const T*)t, (const T*)r, (const T*)t,
                     (const T*)t, N, N, nchunk, part);
  hipLaunchKernelGGL(bi_omega_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, S);
  hipLaunchKernelGGL(bi_x_kernel<T>, gc, dim3(AB), 0, st, x, sp->ld, r, (const T*)rt, (const T*)ph, (const T*)sh,
                     (const T*)t, N, nchunk, (const ColState<D>*)S, part);
  hipLaunchKernelGGL(bi_r_finalize<D>, dim3(ncol), dim3(AB), 0, st, (const D*)part, nchunk, sp->tol, S);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

static int stern_iterations(cyx_stern_plan_s* sp, int iters) {
  int rc = CYX_OK;
  for (int i = 0; i < iters && rc == CYX_OK; ++i) {
    switch (sp->dtype) {
      case CYX_F64: rc = stern_iteration<double>(sp); break;
      case CYX_F32: rc = stern_iteration<float>(sp); break;
      case CYX_C128: rc = stern_iteration<cplx<double>>(sp); break;
      default: rc = stern_iteration<cplx<float>>(sp); break;
    }
  }
  return rc;
}

int stern_plan_enqueue(cyx_stern_plan_s* sp, hipStream_t st, int iters) {
  cyx_ctx_s* c = sp->h->ctx;
  const hipStream_t saved = c->stream;
  const bool timing = c->timing;
  c->stream = st;
  c->timing = false;
  const int rc = stern_iterations(sp, iters);
  c->stream = saved;
  c->timing = timing;
  return rc;
}

cyx_ctx_s* stern_plan_context(const cyx_stern_plan_s* sp) { return sp->h->ctx; }
bool stern_plan_spectral(const cyx_stern_plan_s* sp) { return sp->spectral; }

static void stern_plan_drop_graph(cyx_stern_plan_s* sp) {
  if (sp->exec) (void)hipGraphExecDestroy(sp->exec);
  if (sp->graph) (void)hipGraphDestroy(sp->graph);
  sp->exec = nullptr;
  sp->graph = nullptr;
  sp->graph_iters = 0;
}

// replay `iters` iterations from the plan's graph (captured on first use of
// that count)
static int stern_plan_run(cyx_stern_plan_s* sp, int iters) {
  cyx_ctx_s* c = sp->h->ctx;
  if (sp->spectral) {  // rocFFT transforms are launched directly, not captured
    PhaseTimer tm(c, PH_STERN);
    return stern_iterations(sp, iters);
  }
  if (!sp->exec || sp->graph_iters != iters) {
    if (sp->exec) CYX_HIP(hipStreamSynchronize(c->stream));
    stern_plan_drop_graph(sp);
    // capture on a stream of the plan's own (capture is not allowed on the
    // null stream, which the context may use)
    if (!sp->cap) CYX_HIP(hipStreamCreateWithFlags(&sp->cap, hipStreamNonBlocking));
    CYX_HIP(hipStreamBeginCapture(sp->cap, hipStreamCaptureModeThreadLocal));
    const int rc = stern_plan_enqueue(sp, sp->cap, iters);
    const hipError_t e = hipStreamEndCapture(sp->cap, &sp->graph);
    if (rc != CYX_OK) {
      if (sp->graph) (void)hipGraphDestroy(sp->graph);
      sp->graph = nullptr;
      return rc;
    }
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return CYX_ERR_HIP;
    }
    CYX_HIP(hipGraphInstantiate(&sp->exec, sp->graph, nullptr, nullptr, 0));
    sp->graph_iters = iters;
  }
  PhaseTimer tm(c, PH_STERN);
  CYX_HIP(hipGraphLaunch(sp->exec, c->stream));
  return CYX_OK;
}

template <class D>
static int stern_status_impl(cyx_stern_plan_s* sp, int* n_active, cyx_stern_info* info, int32_t* col_iters,
                             double* col_relres, int* breakdown) {
  hipStream_t st = sp->h->ctx->stream;
  const int ncol = sp->ncol;
  std::vector<ColState<D>> hs(ncol);
  CYX_HIP(hipMemcpyAsync(hs.data(), sp->L.state, ncol * sizeof(ColState<D>), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  int nconv = 0, maxits = 0, brk = 0, act = 0;
  double maxrel = 0.0;
  for (int j = 0; j < ncol; ++j) {
    const double rel = hs[j].bnorm2 > 0 ? std::sqrt(hs[j].rnorm2 / hs[j].bnorm2) : 0.0;
    // a half-step exit leaves r = s, whose norm met the tolerance
    const bool ok = hs[j].bnorm2 == 0.0 || hs[j].half || rel <= sp->tol;
    act += hs[j].active != 0;
    nconv += ok;
    brk += hs[j].flag == 4 && !ok;
    maxits = std::max(maxits, hs[j].iters);
    maxrel = std::max(maxrel, rel);
    if (col_iters) col_iters[j] = hs[j].iters;
    if (col_relres) col_relres[j] = rel;
  }
  if (n_active) *n_active = act;
  if (info) {
    info->max_iterations = maxits;
    info->n_converged = nconv;
    info->max_relres = maxrel;
  }
  if (breakdown) *breakdown = brk;
  return CYX_OK;
}

static int stern_status(cyx_stern_plan_s* sp, int* n_active, cyx_stern_info* info, int32_t* col_iters,
                        double* col_relres, int* breakdown) {
  if (sp->dtype == CYX_C128 || sp->dtype == CYX_C64)
    return stern_status_impl<cplx<double>>(sp, n_active, info, col_iters, col_relres, breakdown);
  return stern_status_impl<double>(sp, n_active, info, col_iters, col_relres, breakdown);
}

// active-column count only (4 bytes back instead of the column states)
template <class D>
static int stern_active_impl(cyx_stern_plan_s* sp, int* active) {
  hipStream_t st = sp->h->ctx->stream;
  hipLaunchKernelGGL(count_active_kernel<D>, dim3(1), dim3(AB), 0, st, (const ColState<D>*)sp->L.state, sp->ncol,
                     sp->L.count);
  CYX_HIP(hipMemcpyAsync(active, sp->L.count, sizeof(int), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  return CYX_OK;
}

static int stern_active(cyx_stern_plan_s* sp, int* active) {
  if (sp->dtype == CYX_C128 || sp->dtype == CYX_C64) return stern_active_impl<cplx<double>>(sp, active);
  return stern_active_impl<double>(sp, active);
}

static size_t stern_plan_layout(cyx_stern_plan_s* sp, char* base) {
  switch (sp->dtype) {
    case CYX_F64: return stern_layout<double>(sp->h, sp->dtype, sp->ncol, sp->nv, sp->nchunk, true, true, base, sp->L);
    case CYX_F32: return stern_layout<float>(sp->h, sp->dtype, sp->ncol, sp->nv, sp->nchunk, true, true, base, sp->L);
    case CYX_C128:
      return stern_layout<cplx<double>>(sp->h, sp->dtype, sp->ncol, sp->nv, sp->nchunk, true, true, base, sp->L);
    default:
      return stern_layout<cplx<float>>(sp->h, sp->dtype, sp->ncol, sp->nv, sp->nchunk, true, true, base, sp->L);
  }
}

static int stern_plan_init(cyx_stern_plan_s* sp, const double* lam, int use_guess) {
  switch (sp->dtype) {
    case CYX_F64: return stern_init<double>(sp, lam, use_guess);
    case CYX_F32: return stern_init<float>(sp, lam, use_guess);
    case CYX_C128: return stern_init<cplx<double>>(sp, lam, use_guess);
    default: return stern_init<cplx<float>>(sp, lam, use_guess);
  }
}

// fills the plan's fields and workspace (own: hipMalloc, else the handle's
// Sternheimer buffer) and computes the initial residuals
static int stern_plan_setup(cyx_stern_plan_s* sp, cyx_ham_s* h, int dtype, const void* B, void* X, int ncol,
                            int64_t ld, const double* lam, const void* PV, int64_t ldv, int nv, int use_guess,
                            double tol, int precond, bool own) {
  sp->h = h;
  sp->dtype = dtype;
  sp->ncol = ncol;
  sp->nv = nv;
  sp->B = B;
  sp->X = X;
  sp->ld = ld;
  sp->PV = PV;
  sp->ldv = ldv;
  sp->tol = tol;
  CYX_CHECK(precond >= 0 && precond <= 2, "precond must be 0 (none), 1 (Jacobi) or 2 (spectral)");
  sp->precond = precond == 1;
  sp->spectral = precond == 2;
  if (sp->spectral) {
    CYX_CHECK(dtype == CYX_F64 || dtype == CYX_C128, "the spectral preconditioner needs FP64 blocks");
    if (!h->spec) {
      const int rc = spectral_create(h->g, &h->spec);
      if (rc != CYX_OK) return rc;
    }
  }
  sp->nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(128, ceil_div(h->N(), 2048)));
  const size_t bytes = stern_plan_layout(sp, nullptr);
  if (own) {
    CYX_CHECK(hipMalloc(&sp->ws, bytes) == hipSuccess, "Sternheimer plan workspace allocation failed (%zu bytes)",
              bytes);
    sp->owns_ws = true;
  } else {
    CYX_CHECK(ensure(h->stern, bytes) == CYX_OK, "Sternheimer workspace allocation failed");
    sp->ws = h->stern.p;
  }
  stern_plan_layout(sp, (char*)sp->ws);
  return stern_plan_init(sp, lam, use_guess);
}

static void stern_plan_release(cyx_stern_plan_s* sp) {
  // a replay may still run
  if (sp->exec) (void)hipStreamSynchronize(sp->h->ctx->stream);
  stern_plan_drop_graph(sp);
  if (sp->cap) (void)hipStreamDestroy(sp->cap);
  sp->cap = nullptr;
  if (sp->owns_ws) (void)hipFree(sp->ws);
  sp->ws = nullptr;
}

// iterations in graph replays of `poll` iterations, the active count read
// after each (the previous host-polled loop's cadence)
static int stern_solve_impl(cyx_ham_s* h, int dtype, const void* B, void* X, int ncol, int64_t ld, const double* lam,
                            const void* PV, int64_t ldv, int nv, int use_guess, double tol, int maxit, int precond,
                            cyx_stern_info* info, int32_t* col_iters, double* col_relres) {
  cyx_stern_plan_s sp;
  int rc = stern_plan_setup(&sp, h, dtype, B, X, ncol, ld, lam, PV, ldv, nv, use_guess, tol, precond, false);
  int active = ncol, it = 0;
  const int poll = 4;
  if (rc == CYX_OK) rc = stern_active(&sp, &active);
  while (rc == CYX_OK && it < maxit && active > 0) {
    const int k = std::min(poll, maxit - it);
    // a short final round runs without a graph of its own
    rc = k == poll ? stern_plan_run(&sp, k) : stern_plan_enqueue(&sp, h->ctx->stream, k);
    it += k;
    if (rc == CYX_OK) rc = stern_active(&sp, &active);
  }
  int breakdown = 0;
  cyx_stern_info inf{};
  if (rc == CYX_OK) rc = stern_status(&sp, nullptr, &inf, col_iters, col_relres, &breakdown);
  stern_plan_release(&sp);
  if (rc != CYX_OK) return rc;
  if (info) *info = inf;
  if (breakdown) {
    set_error("BiCGSTAB broke down (rho or omega zero) on %d of %d columns", breakdown, ncol);
    return CYX_ERR_NOCONV;
  }
  if (inf.n_converged < ncol) {
    set_error("BiCGSTAB reached %d iterations with %d of %d columns above tol %g", maxit, ncol - inf.n_converged,
              ncol, tol);
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
  CYX_CHECK(ncol >= 1 && ncol <= 65535 && ld >= h->N(), "bad block shape (ncol=%d ld=%lld N=%lld)", ncol,
            (long long)ld, (long long)h->N());
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
  return stern_solve_impl(h, dtype, B, X, ncol, ld, lam, psi_v, ldv, nv, use_guess, tol, max_iter, precond, info,
                          col_iters, col_relres);
}

int cyx_sternheimer_plan_create(cyx_ham_t h, int dtype, const void* B, void* X, int ncol, int64_t ld,
                                const double* lam, const void* psi_v, int64_t ldv, int nv, int use_guess, double tol,
                                int precond, cyx_stern_plan_t* out) {
  int rc;
  CYX_CHECK(out, "null argument");
  *out = nullptr;
  if ((rc = stern_check(h, dtype, ncol, ld, lam, psi_v, ldv, nv)) != CYX_OK) return rc;
  CYX_CHECK(B && X && B != X, "null or aliased argument");
  CYX_CHECK(tol > 0, "bad tolerance");
  cyx_stern_plan_s* sp = new cyx_stern_plan_s();
  rc = stern_plan_setup(sp, h, dtype, B, X, ncol, ld, lam, psi_v, ldv, nv, use_guess, tol, precond, true);
  if (rc != CYX_OK) {
    (void)hipStreamSynchronize(h->ctx->stream);
    stern_plan_release(sp);
    delete sp;
    return rc;
  }
  *out = sp;
  return CYX_OK;
}

int cyx_sternheimer_plan_iterate(cyx_stern_plan_t sp, int iters) {
  CYX_CHECK(sp && iters >= 1, "bad argument");
  return stern_plan_run(sp, iters);
}

int cyx_sternheimer_plan_status(cyx_stern_plan_t sp, int* n_active, cyx_stern_info* info, int32_t* col_iters,
                                double* col_relres) {
  CYX_CHECK(sp, "null argument");
  return stern_status(sp, n_active, info, col_iters, col_relres, nullptr);
}

int cyx_sternheimer_plan_destroy(cyx_stern_plan_t sp) {
  if (!sp) return CYX_OK;
  (void)hipDeviceSynchronize();  // the context may already be gone
  stern_plan_release(sp);
  delete sp;
  return CYX_OK;
}

int cyx_sternheimer_rhs(cyx_ham_t dham, int dtype, const void* psi, int64_t ldpsi, int nv, const double* coeff,
                        void* B, int64_t ldb, double* dlambda) {
  CYX_CHECK(dham && psi && coeff && B && psi != B, "null or aliased argument");
  CYX_CHECK(dtype == CYX_F64 || dtype == CYX_C128 || dtype == CYX_F32 || dtype == CYX_C64, "bad dtype %d", dtype);
  CYX_CHECK(!dham->bloch || dtype == CYX_C128 || dtype == CYX_C64, "a Bloch channel needs a complex dtype");
  CYX_CHECK(nv >= 1 && nv <= 65535 && ldpsi >= dham->N() && ldb >= dham->N(), "bad block shape (nv=%d)", nv);
  switch (dtype) {
    case CYX_F64: return stern_rhs_impl<double>(dham, dtype, psi, ldpsi, nv, coeff, B, ldb, dlambda);
    case CYX_F32: return stern_rhs_impl<float>(dham, dtype, psi, ldpsi, nv, coeff, B, ldb, dlambda);
    case CYX_C128: return stern_rhs_impl<cplx<double>>(dham, dtype, psi, ldpsi, nv, coeff, B, ldb, dlambda);
    default: return stern_rhs_impl<cplx<float>>(dham, dtype, psi, ldpsi, nv, coeff, B, ldb, dlambda);
  }
}

}  // extern "C"
