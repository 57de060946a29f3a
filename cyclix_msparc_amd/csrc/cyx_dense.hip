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

// ---- The same reduction spread over G workgroups (one per CU) with the
// matrix resident in LDS.  Column j lives in workgroup j % G (cyclic, so the
// shrinking trailing matrix stays balanced), Hermitian in full, built from
// the lower triangle.  Step k needs one exchange between workgroups: each
// publishes its partial of A22 v_k (rows > k) and the owner of column k+1
// publishes that column; after a grid barrier every workgroup sums the
// partials in the same order, so all of them compute bit-identical w_k, the
// updated column k+1 and the next reflector v_{k+1} without further
// communication.  Each workgroup then applies the rank-2 update of step k to
// its own columns and, in the same pass over LDS, accumulates its partial of
// A22 v_{k+1}.  Partials and the published column are double-buffered by the
// parity of k: buffer k & 1 is rewritten in step k + 1, after every workgroup
// has passed barrier k + 1 and so has finished reading it.  The launch is
// cooperative (all G workgroups resident); every barrier wait is bounded and
// sets sync[1] when it gives up, after which all workgroups return.
constexpr int TRD2_NT = 256;
constexpr size_t TRD2_LDS = 159 * 1024;  // dynamic LDS budget of one workgroup
constexpr int TRD2_GMAX = 32;
constexpr unsigned TRD2_SPIN = 1u << 22;

// grid barrier number `target / G`: every workgroup adds one to sync[0]
__device__ __forceinline__ bool trd_grid_barrier(unsigned* sync, unsigned target, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int good = 1;
    unsigned polls = 0;
    while (__hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++polls > TRD2_SPIN) {
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
    if (!good) __hip_atomic_fetch_or(sync + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *s_ok = good;
  }
  __syncthreads();
  return *s_ok != 0;
}

// Reflector of x = ck[r0 .. n) (alpha = ck[r0]), zlarfg: v = (1, x[1..] /
// (alpha - beta)) into vout[r0 .. n), tau and beta into *tau_out, *beta_out.
template <class D>
__device__ __forceinline__ void trd_reflector(const D* ck, int r0, int n, D* vout, D* tau_out, double* beta_out,
                                              D* s_scale, double* red) {
  const int t = threadIdx.x;
  double q[1] = {0.0};
  for (int i = r0 + 1 + t; i < n; i += TRD2_NT) q[0] += absq(ck[i]);
  const double xn2 = block_sums<1, TRD2_NT>(q, red);
  if (t == 0) {
    const D alpha = ck[r0];
    const double ar = re_of(alpha), ai = im_of(alpha);
    if (xn2 == 0.0 && ai == 0.0) {
      *tau_out = make_d<D>(0.0, 0.0);
      *s_scale = make_d<D>(0.0, 0.0);
      *beta_out = ar;
    } else {
      const double beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
      *tau_out = make_d<D>((beta - ar) / beta, -ai / beta);
      const double zr = ar - beta, zi = ai, den = zr * zr + zi * zi;
      *s_scale = make_d<D>(zr / den, -zi / den);
      *beta_out = beta;
    }
  }
  __syncthreads();
  const D sc = *s_scale;
  for (int i = r0 + t; i < n; i += TRD2_NT) vout[i] = i == r0 ? make_d<D>(1.0, 0.0) : cmul(sc, ck[i]);
  __syncthreads();
}

template <class D>
__global__ void __launch_bounds__(TRD2_NT)
    hetrd_dist_kernel(D* __restrict__ A, int n, int G, int cpw, double* __restrict__ d, double* __restrict__ e,
                      D* __restrict__ tau, D* __restrict__ pbuf, D* __restrict__ cbuf, unsigned* __restrict__ sync) {
  extern __shared__ __align__(16) unsigned char trd2_raw[];
  D* slab = reinterpret_cast<D*>(trd2_raw);  // [cpw][n]: local column lc is column lc * G + g
  D* vv = slab + (size_t)cpw * n;            // [2][n] reflectors v_k and v_{k+1}
  D* ww = vv + 2 * n;                        // [n] w_k
  D* ck = ww + n;                            // [n] column k + 1
  __shared__ double red[(TRD2_NT / 64) * 2];
  __shared__ double s_dot[2], s_beta[2];
  __shared__ D s_tau[2], s_scale;
  __shared__ int s_ok;
  const int t = threadIdx.x, g = blockIdx.x;
  const int lce = g < n ? min(cpw, (n - 1 - g) / G + 1) : 0;  // own columns lc * G + g < n
  for (int lc = 0; lc < lce; ++lc) {
    const int j = lc * G + g;
    for (int i = t; i < n; i += TRD2_NT) {
      D a = i >= j ? A[(int64_t)j * n + i] : conjg(A[(int64_t)i * n + j]);
      if (i == j) a = make_d<D>(re_of(a), 0.0);
      slab[(size_t)lc * n + i] = a;
    }
  }
  for (int i = t; i < n; i += TRD2_NT) ck[i] = A[i];
  __syncthreads();
  if (g == 0 && t == 0) d[0] = re_of(ck[0]);
  if (n < 2) return;
  int cur = 0;
  trd_reflector(ck, 1, n, vv, &s_tau[0], &s_beta[0], &s_scale, red);
  {  // partial of A22 v_0 and column 1
    const int lc0 = g >= 1 ? 0 : 1;
    for (int i = 1 + t; i < n; i += TRD2_NT) {
      D acc = make_d<D>(0.0, 0.0);
      for (int lc = lc0; lc < lce; ++lc) {
        const int j = lc * G + g;
        acc = cfma(slab[(size_t)lc * n + i], vv[j], acc);
      }
      pbuf[(size_t)g * n + i] = acc;
      if (1 % G == g) cbuf[i] = slab[(size_t)(1 / G) * n + i];
    }
  }
  for (int k = 0; k <= n - 2; ++k) {
    const int buf = k & 1, nxt = cur ^ 1;
    if (!trd_grid_barrier(sync, (unsigned)(k + 1) * (unsigned)G, &s_ok)) return;
    const D* vc = vv + (size_t)cur * n;
    const D tk = s_tau[cur];
    // w = tau A22 v (the partials in workgroup order), w^H v, column k + 1
    double dv[2] = {0.0, 0.0};
    for (int i = k + 1 + t; i < n; i += TRD2_NT) {
      const D* pb = pbuf + (size_t)buf * G * n + i;
      const D cv = cbuf[(size_t)buf * n + i];
      // groups of 8 loads in flight, summed in workgroup order
      D sacc = make_d<D>(0.0, 0.0);
      for (int q0 = 0; q0 < G; q0 += 8) {
        D part[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) part[u] = q0 + u < G ? pb[(size_t)(q0 + u) * n] : make_d<D>(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < 8; ++u) sacc = add(sacc, part[u]);
      }
      const D wi = cmul(tk, sacc);
      ww[i] = wi;
      const D c = mul_cj(vc[i], wi);
      dv[0] += re_of(c);
      dv[1] += im_of(c);
      ck[i] = cv;
    }
    if (g == 0) {  // LAPACK storage of step k: e, tau, beta on the subdiagonal, v below it
      if (t == 0) {
        e[k] = s_beta[cur];
        tau[k] = tk;
      }
      for (int i = k + 1 + t; i < n; i += TRD2_NT)
        A[(int64_t)k * n + i] = i == k + 1 ? make_d<D>(s_beta[cur], 0.0) : vc[i];
    }
    const double dsum = block_sums<2, TRD2_NT>(dv, red);
    if (t < 2) s_dot[t] = dsum;
    __syncthreads();
    {
      const D a2 = rmul(-0.5, cmul(tk, make_d<D>(s_dot[0], s_dot[1])));
      for (int i = k + 1 + t; i < n; i += TRD2_NT) ww[i] = cfma(a2, vc[i], ww[i]);
    }
    __syncthreads();
    {  // column k + 1 after step k
      const D wk = ww[k + 1], vk = vc[k + 1];
      for (int i = k + 1 + t; i < n; i += TRD2_NT) {
        D a = sub(ck[i], add(mul_cj(vc[i], wk), mul_cj(ww[i], vk)));
        if (i == k + 1) a = make_d<D>(re_of(a), 0.0);
        ck[i] = a;
      }
    }
    __syncthreads();
    if (g == 0 && t == 0) d[k + 1] = re_of(ck[k + 1]);
    if (k == n - 2) break;
    trd_reflector(ck, k + 2, n, vv + (size_t)nxt * n, &s_tau[nxt], &s_beta[nxt], &s_scale, red);
    // rank-2 update of the own columns j >= k + 2 (rows >= k + 2) and the
    // partial of A22 v_{k+1}; the owner of column k + 2 publishes it
    const D* vn = vv + (size_t)nxt * n;
    const int lc0 = k + 2 > g ? (k + 2 - g + G - 1) / G : 0;
    const bool pub = (k + 2) % G == g;
    const int lpub = (k + 2) / G;
    D* pout = pbuf + ((size_t)(buf ^ 1) * G + g) * n;
    for (int i = k + 2 + t; i < n; i += TRD2_NT) {
      const D vi = vc[i], wi = ww[i];
      D acc = make_d<D>(0.0, 0.0);
#pragma unroll 4
      for (int lc = lc0; lc < lce; ++lc) {
        const int j = lc * G + g;
        D* a = slab + (size_t)lc * n + i;
        D na = sub(*a, add(mul_cj(vi, ww[j]), mul_cj(wi, vc[j])));
        if (i == j) na = make_d<D>(re_of(na), 0.0);
        *a = na;
        acc = cfma(na, vn[j], acc);
      }
      pout[i] = acc;
      if (pub) cbuf[(size_t)(buf ^ 1) * n + i] = slab[(size_t)lpub * n + i];
    }
    cur = nxt;
  }
}

// workgroups of the distributed reduction for n (0: does not fit)
int hetrd_dist_groups(int n, bool cx) {
  const size_t esz = cx ? 16 : 8, col = (size_t)n * esz, fixed = 4 * col;
  if (n < 1 || fixed >= TRD2_LDS) return 0;
  const size_t cmax = std::min<size_t>((TRD2_LDS - fixed) / col, (size_t)n);
  if (cmax < 1) return 0;
  const int G = (int)((n + cmax - 1) / cmax);
  return G <= TRD2_GMAX ? G : 0;
}

size_t hetrd_dist_ws_bytes(int n, bool cx) {
  const int G = hetrd_dist_groups(n, cx);
  return G ? (2 * (size_t)G + 2) * n * (cx ? 16 : 8) + 256 : 0;
}

// ws: hetrd_dist_ws_bytes(n, cx) bytes, 256-byte aligned; its last 256 bytes
// hold the barrier counter and the failure flag (sync[1], nonzero when a
// barrier wait gave up), cleared here before the launch.  Not for use inside
// stream capture (the clear is a memset).
int hetrd_dist(hipStream_t st, bool cx, void* A, int n, double* d, double* e, void* tau, void* ws,
               unsigned** flag) {
  const int G = hetrd_dist_groups(n, cx);
  CYX_CHECK(G > 0, "distributed tridiagonal reduction of n = %d does not fit %d workgroups", n, TRD2_GMAX);
  const size_t esz = cx ? 16 : 8;
  const int cpw = (n + G - 1) / G;
  const size_t shm = (4 + (size_t)cpw) * n * esz;
  char* w = (char*)ws;
  void* pbuf = w;
  void* cbuf = w + 2 * (size_t)G * n * esz;
  unsigned* sync = (unsigned*)(w + (2 * (size_t)G + 2) * n * esz);
  *flag = sync + 1;
  CYX_HIP(hipMemsetAsync(sync, 0, 2 * sizeof(unsigned), st));
  const void* fn = cx ? (const void*)&hetrd_dist_kernel<cplx<double>> : (const void*)&hetrd_dist_kernel<double>;
  CYX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  void* args[] = {&A, &n, (void*)&G, (void*)&cpw, &d, &e, &tau, &pbuf, &cbuf, &sync};
  if (hipLaunchCooperativeKernel(fn, dim3(G), dim3(TRD2_NT), args, (unsigned)shm, st) != hipSuccess) {
    // the runtime refused a co-resident launch: the one-workgroup reduction
    // gives the same factorisation (its failure word stays clear)
    (void)hipGetLastError();
    if (hetrd_lds_bytes(n, cx) <= 64 * 1024) return hetrd(st, cx, A, n, d, e, tau);
    set_error("tridiagonal reduction of n = %d: cooperative launch refused", n);
    return CYX_ERR_HIP;
  }
  return CYX_OK;
}

}  // namespace cyx
