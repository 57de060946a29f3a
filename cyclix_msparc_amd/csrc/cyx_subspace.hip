// Subspace kernels of the Rayleigh-Ritz step (src/eigSolver.m:99-124):
// weighted Gram/projection M = A^H diag(W) B, rotation A*Q, column
// W-normalisation.  All products accumulate in FP64 (FP32 inputs are
// widened exactly, so the Gram of an FP32 block carries FP64 accuracy); the
// Gram and the rotation run on the FP64 matrix cores.
#include "cyx_internal.hpp"

namespace cyx {

// Both products run on the FP64 matrix cores, v_mfma_f64_16x16x4_f64: lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and receives D[4t + l/16][l%16] in
// element t of its accumulator.  FP32 inputs are widened exactly on load.
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---- Gram M = A^H diag(W) B.  A wave owns a 32 x 32 output tile (2 x 2
// MFMA tiles) over a k sub-range; the 4 waves of a block take consecutive
// sub-ranges of one tile and sum them in LDS, one partial per block.
// The k slots of an MFMA may hold any points as long as A and B agree, so
// lane (g, c) takes points n0 + SP g + s (s < SP) of column c: its loads run
// along a column, the 4 lane groups of a column reading one 128-byte line per
// step.  The k loop is software-pipelined: the loads of step k+1 are issued
// before the MFMAs of step k (scheduling barriers keep that order), which
// the compiler's own schedule does not do.  Hermitian products compute the
// tiles with ta <= tb only.
constexpr int GS = 32;  // output tile of a wave
template <class T>
constexpr int gram_sp() {  // points per lane per k step
  return sizeof(T) >= 16 ? 2 : (sizeof(T) >= 8 ? 4 : 8);
}

__device__ inline void gram_tile_of(int tp, bool herm, int nta, int& ta, int& tb) {
  if (herm) {
    tb = 0;
    while ((tb + 1) * (tb + 2) / 2 <= tp) ++tb;
    ta = tp - tb * (tb + 1) / 2;
  } else {
    ta = tp % nta;
    tb = tp / nta;
  }
}

template <class T, int SP>
struct GramStep {
  T a[2][SP], b[2][SP];
  double w[SP];
};

// a lane's SP consecutive points of one column (32 bytes) as two 16-byte loads
typedef unsigned v4u __attribute__((ext_vector_type(4)));
template <class T, int SP>
__device__ __forceinline__ void gram_run(T (&dst)[SP], const T* __restrict__ p) {
  static_assert(sizeof(T) * SP == 32, "a run is 32 bytes");
  const v4u* q = reinterpret_cast<const v4u*>(p);
  const v4u u0 = q[0], u1 = q[1];
  const unsigned w[8] = {u0[0], u0[1], u0[2], u0[3], u1[0], u1[1], u1[2], u1[3]};
  __builtin_memcpy(dst, w, 32);
}

// vec: the leading dimensions keep every lane's run 16-byte aligned
template <class T, int SP>
__device__ __forceinline__ void gram_load(GramStep<T, SP>& S, const T* __restrict__ A, int64_t lda, int na,
                                          const T* __restrict__ B, int64_t ldb, int nb,
                                          const double* __restrict__ W, int64_t n0, int64_t k1, const int (&ca)[2],
                                          const int (&cb)[2], int g, bool vec) {
  const int64_t nl = n0 + SP * g;
  if (vec && nl + SP <= k1) {
#pragma unroll
    for (int s = 0; s < SP; ++s) S.w[s] = W[nl + s];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (ca[i] < na) {
        gram_run(S.a[i], A + (int64_t)ca[i] * lda + nl);
      } else {
#pragma unroll
        for (int s = 0; s < SP; ++s) S.a[i][s] = zero<T>();
      }
      if (cb[i] < nb) {
        gram_run(S.b[i], B + (int64_t)cb[i] * ldb + nl);
      } else {
#pragma unroll
        for (int s = 0; s < SP; ++s) S.b[i][s] = zero<T>();
      }
    }
    return;
  }
#pragma unroll
  for (int s = 0; s < SP; ++s) {
    const int64_t n = n0 + SP * g + s;
    const bool ok = n < k1;
    S.w[s] = ok ? W[n] : 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      S.a[i][s] = (ok && ca[i] < na) ? A[(int64_t)ca[i] * lda + n] : zero<T>();
      S.b[i][s] = (ok && cb[i] < nb) ? B[(int64_t)cb[i] * ldb + n] : zero<T>();
    }
  }
}

template <class T, int SP, int NC>
__device__ __forceinline__ void gram_mfma(const GramStep<T, SP>& S, v4d (&acc)[NC][2][2]) {
#pragma unroll
  for (int s = 0; s < SP; ++s) {
    if constexpr (traits<T>::is_complex) {
      // conj(a) w b: re += ar br + ai bi, im += ar bi - ai br
      cplx<double> av[2];
      double br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = widen(S.a[i][s]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const cplx<double> bv = widen(S.b[j][s]);
        br[j] = S.w[s] * bv.x;
        bi[j] = S.w[s] * bv.y;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[0][i][j] = mfma64(av[i].x, br[j], acc[0][i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[1][i][j] = mfma64(av[i].x, bi[j], acc[1][i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[0][i][j] = mfma64(av[i].y, bi[j], acc[0][i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[1][i][j] = mfma64(-av[i].y, br[j], acc[1][i][j]);
    } else {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[0][i][j] = mfma64(widen(S.a[i][s]), S.w[s] * widen(S.b[j][s]), acc[0][i][j]);
    }
  }
}

template <class T>
__global__ void __launch_bounds__(256, 2) gram_mfma_kernel(const T* __restrict__ A, int64_t lda, int na,
                                                        const T* __restrict__ B, int64_t ldb, int nb,
                                                        const double* __restrict__ W, int64_t N, int64_t chunk,
                                                        int herm, int nta, int vec,
                                                        typename acc_of<T>::type* __restrict__ part) {
  using D = typename acc_of<T>::type;
  constexpr bool CX = traits<T>::is_complex;
  constexpr int NC = CX ? 2 : 1;  // accumulator sets: (re, im) or one
  constexpr int SP = gram_sp<T>();
  __shared__ double red[4][NC][2][2][4][64];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  int ta, tb;
  gram_tile_of(blockIdx.x, herm != 0, nta, ta, tb);
  const int a0 = ta * GS, b0 = tb * GS;
  const int64_t k0 = ((int64_t)blockIdx.y * 4 + wv) * chunk, k1 = min(N, k0 + chunk);
  v4d acc[NC][2][2];
#pragma unroll
  for (int q = 0; q < NC; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[q][i][j] = v4d{0, 0, 0, 0};
  const int ca[2] = {a0 + c, a0 + 16 + c}, cb[2] = {b0 + c, b0 + 16 + c};
  if (k0 < k1) {
    GramStep<T, SP> S0, S1;
    gram_load(S0, A, lda, na, B, ldb, nb, W, k0, k1, ca, cb, g, vec != 0);
    for (int64_t n0 = k0;;) {
      const bool more = n0 + 4 * SP < k1;
      if (more) gram_load(S1, A, lda, na, B, ldb, nb, W, n0 + 4 * SP, k1, ca, cb, g, vec != 0);
      __builtin_amdgcn_sched_barrier(0);
      gram_mfma(S0, acc);
      __builtin_amdgcn_sched_barrier(0);
      if (!more) break;
      n0 += 4 * SP;
      const bool more2 = n0 + 4 * SP < k1;
      if (more2) gram_load(S0, A, lda, na, B, ldb, nb, W, n0 + 4 * SP, k1, ca, cb, g, vec != 0);
      __builtin_amdgcn_sched_barrier(0);
      gram_mfma(S1, acc);
      __builtin_amdgcn_sched_barrier(0);
      if (!more2) break;
      n0 += 4 * SP;
    }
  }
#pragma unroll
  for (int q = 0; q < NC; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int t = 0; t < 4; ++t) red[wv][q][i][j][t][l] = acc[q][i][j][t];
  __syncthreads();
  // the 32 x 32 tile: element e -> (i, j, t, lane) of the MFMA layout, the 4 waves summed in order
  D* out = part + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * (GS * GS);
  for (int e = threadIdx.x; e < GS * GS; e += 256) {
    const int ll = e & 63, t = (e >> 6) & 3, j = (e >> 8) & 1, i = (e >> 9) & 1;
    const int al = 16 * i + 4 * t + (ll >> 4), bl = 16 * j + (ll & 15);
    D v{};
    if constexpr (CX) {
      double sr = 0, si = 0;
      for (int w2 = 0; w2 < 4; ++w2) {
        sr += red[w2][0][i][j][t][ll];
        si += red[w2][1][i][j][t][ll];
      }
      v = D{sr, si};
    } else {
      double sr = 0;
      for (int w2 = 0; w2 < 4; ++w2) sr += red[w2][0][i][j][t][ll];
      v = sr;
    }
    out[bl * GS + al] = v;
  }
}

// M (na x nb column-major) = fixed-order sum of the block partials; Hermitian
// products keep the upper triangle and mirror it (conjugated) into the lower
// A block takes GRE elements of one tile; its 256 threads are GRE elements x
// GRL chunk lanes, lane l summing the chunks k = l mod GRL in order, and the
// lanes' sums are added in lane order (a fixed order, so results are
// reproducible), so that products with few tiles and many chunks still
// spread over GS*GS/GRE blocks per tile.
constexpr int GRE = 64, GRL = 4;

template <class D>
__global__ void __launch_bounds__(GRE * GRL) gram_reduce_kernel(const D* __restrict__ part, int nchb, int herm,
                                                                int nta, int na, int nb, D* __restrict__ M) {
  __shared__ D red[GRL][GRE];
  const int tp = blockIdx.x;
  const int el = threadIdx.x % GRE, lane = threadIdx.x / GRE;
  const int e = blockIdx.y * GRE + el;
  D s{};
  for (int k = lane; k < nchb; k += GRL) s = add(s, part[((int64_t)tp * nchb + k) * (GS * GS) + e]);
  red[lane][el] = s;
  __syncthreads();
  if (lane != 0) return;
#pragma unroll
  for (int l = 1; l < GRL; ++l) s = add(s, red[l][el]);
  int ta, tb;
  gram_tile_of(tp, herm != 0, nta, ta, tb);
  const int a = ta * GS + (e % GS), b = tb * GS + (e / GS);
  if (a < na && b < nb) {
    if (!herm) {
      M[(int64_t)b * na + a] = s;
    } else if (a <= b) {  // diagonal tiles hold both triangles: keep the upper one
      M[(int64_t)b * na + a] = s;
      if (a != b) M[(int64_t)a * na + b] = conjg(s);
    }
  }
}

// ---- rotation out = A Q, computed as out^T = Q^T A^T so that lane (g, c)
// holds out[n0 + c][b0 + 4t + g]: 16 consecutive points of one column per
// store.  A block of RW waves covers 16 RW points and all ns columns in
// passes of 64 (4 MFMA tiles); Q[a chunk][64 columns] is staged in LDS.
constexpr int RW = 8;    // waves per rotation block
constexpr int RAC = 16;  // a per LDS chunk

template <class T>
__global__ void __launch_bounds__(RW * 64) rotate_mfma_kernel(const T* __restrict__ A,
                                                              const typename acc_of<T>::type* __restrict__ Q,
                                                              T* __restrict__ out, int64_t N, int64_t ld, int ns) {
  using D = typename acc_of<T>::type;
  constexpr bool CX = traits<T>::is_complex;
  constexpr int NC = CX ? 2 : 1;
  __shared__ D sq[RAC][64 + 1];  // [a][b]
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  const int64_t n = ((int64_t)blockIdx.x * RW + wv) * 16 + c;
  const bool nok = n < N;
  for (int bg = 0; bg < ns; bg += 64) {
    v4d acc[NC][4];
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int bt = 0; bt < 4; ++bt) acc[q][bt] = v4d{0, 0, 0, 0};
    for (int ac = 0; ac < ns; ac += RAC) {
      __syncthreads();
      for (int e = threadIdx.x; e < RAC * 64; e += RW * 64) {
        const int aa = e % RAC, bb = e / RAC;
        sq[aa][bb] = (ac + aa < ns && bg + bb < ns) ? Q[(int64_t)(bg + bb) * ns + ac + aa] : D{};
      }
      __syncthreads();
#pragma unroll
      for (int k4 = 0; k4 < RAC; k4 += 4) {
        const int a = ac + k4 + g;
        const D x = (nok && a < ns) ? widen(A[(int64_t)a * ld + n]) : D{};
#pragma unroll
        for (int bt = 0; bt < 4; ++bt) {
          const D qv = sq[k4 + g][bt * 16 + c];
          if constexpr (CX) {
            // x q: re += xr qr - xi qi, im += xi qr + xr qi
            acc[0][bt] = mfma64(qv.x, x.x, acc[0][bt]);
            acc[1][bt] = mfma64(qv.x, x.y, acc[1][bt]);
            acc[0][bt] = mfma64(-qv.y, x.y, acc[0][bt]);
            acc[1][bt] = mfma64(qv.y, x.x, acc[1][bt]);
          } else {
            acc[0][bt] = mfma64(qv, x, acc[0][bt]);
          }
        }
      }
    }
    if (nok) {
#pragma unroll
      for (int bt = 0; bt < 4; ++bt)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int b = bg + 16 * bt + 4 * t + g;
          if (b < ns) {
            D v;
            if constexpr (CX)
              v = D{acc[0][bt][t], acc[1][bt][t]};
            else
              v = acc[0][bt][t];
            out[(int64_t)b * ld + n] = narrow<T>(v);
          }
        }
    }
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

// res[s] = sqrt(sum W |ha - lam_s a|^2 / sum W |a|^2), FP64 accumulation
template <class T>
__global__ void __launch_bounds__(1024) resnorm_kernel(const T* __restrict__ A, const T* __restrict__ HA,
                                                       const double* __restrict__ lam, const double* __restrict__ W,
                                                       int64_t N, int64_t ld, double* __restrict__ res) {
  __shared__ double red[2][1024];
  const int s = blockIdx.x;
  const T* a = A + (int64_t)s * ld;
  const T* ha = HA + (int64_t)s * ld;
  const double l = lam[s];
  double rr = 0, aa = 0;
  for (int64_t n = threadIdx.x; n < N; n += blockDim.x) {
    const auto x = widen(a[n]);
    rr += W[n] * absq(rfma(-l, x, widen(ha[n])));
    aa += W[n] * absq(x);
  }
  red[0][threadIdx.x] = rr;
  red[1][threadIdx.x] = aa;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) res[s] = sqrt(red[0][0] / red[1][0]);
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
struct GramShape {
  int nta, ntp, nchb;
  int64_t chunk;
  size_t pbytes;
};

template <class T>
static GramShape gram_shape(int64_t N, int na, int nb, bool herm) {
  using D = typename acc_of<T>::type;
  GramShape g;
  g.nta = ceil_div(na, GS);
  const int ntb = ceil_div(nb, GS);
  g.ntp = herm ? g.nta * (g.nta + 1) / 2 : g.nta * ntb;
  // about 8192 waves in all (4 per block; 16384 for a Hermitian product,
  // whose 36 of 64 tiles leave the machine short of waves otherwise:
  // measured 4.3 against 4.46 ms per step for both products, with 4096 or
  // 32768 slower), k sub-ranges in whole k steps
  // (at most 512 chunks: a thin product, whose few tiles would otherwise be
  // cut into thousands of chunks, then moves 16 MB of partials at most)
  int nchb = std::max(1, std::min(512, ceil_div(herm ? 16384 : 8192, 4 * g.ntp)));
  const int kstep = 4 * gram_sp<T>();
  int64_t chunk = ceil_div(N, 4 * (int64_t)nchb);
  chunk = ceil_div(chunk, kstep) * kstep;
  g.nchb = (int)ceil_div(N, 4 * chunk);
  g.chunk = chunk;
  g.pbytes = (size_t)g.ntp * g.nchb * GS * GS * sizeof(D);
  return g;
}

size_t gram_part_bytes(int dtype, int64_t N, int na, int nb, bool herm) {
  herm = herm && na == nb;
  switch (dtype) {
    case CYX_F64:
      return gram_shape<double>(N, na, nb, herm).pbytes;
    case CYX_F32:
      return gram_shape<float>(N, na, nb, herm).pbytes;
    case CYX_C128:
      return gram_shape<cplx<double>>(N, na, nb, herm).pbytes;
    case CYX_C64:
      return gram_shape<cplx<float>>(N, na, nb, herm).pbytes;
  }
  return 0;
}

template <class T>
static int gram_impl(cyx_ham_s* h, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                     const double* W, void* M_out, bool herm, void* part) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  herm = herm && na == nb;
  const GramShape gs = gram_shape<T>(N, na, nb, herm);
  const int nta = gs.nta, ntp = gs.ntp, nchb = gs.nchb;
  const int64_t chunk = gs.chunk;
  if (!part) {
    CYX_CHECK(ensure(h->small, gs.pbytes) == CYX_OK, "partials allocation failed");
    part = h->small.p;
  }
  hipStream_t st = h->ctx->stream;
  PhaseTimer tm(h->ctx, PH_GRAM);
  // 16-byte vector loads when every lane's 32-byte run of a column is aligned
  const auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const int vec = ((lda * sizeof(T)) % 16 == 0 && (ldb * sizeof(T)) % 16 == 0 && al(A) && al(B)) ? 1 : 0;
  hipLaunchKernelGGL(gram_mfma_kernel<T>, dim3(ntp, nchb), dim3(256), 0, st, (const T*)A, lda, na, (const T*)B, ldb,
                     nb, W, N, chunk, herm ? 1 : 0, nta, vec, (D*)part);
  CYX_HIP(hipGetLastError());
  hipLaunchKernelGGL(gram_reduce_kernel<D>, dim3(ntp, GS * GS / GRE), dim3(GRE * GRL), 0, st, (const D*)part, nchb,
                     herm ? 1 : 0, nta, na, nb, (D*)M_out);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int gram_rect(cyx_ham_s* h, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
              const double* W, void* M_out, bool herm, void* part) {
  switch (dtype) {
    case CYX_F64:
      return gram_impl<double>(h, A, lda, na, B, ldb, nb, W, M_out, herm, part);
    case CYX_F32:
      return gram_impl<float>(h, A, lda, na, B, ldb, nb, W, M_out, herm, part);
    case CYX_C128:
      return gram_impl<cplx<double>>(h, A, lda, na, B, ldb, nb, W, M_out, herm, part);
    case CYX_C64:
      return gram_impl<cplx<float>>(h, A, lda, na, B, ldb, nb, W, M_out, herm, part);
  }
  return CYX_ERR_ARG;
}

int gram_w(cyx_ham_s* h, int dtype, const void* A, const void* B, int ns, int64_t ld, void* M_out) {
  // A^H W A is Hermitian by construction; a product of two blocks is not
  return gram_rect(h, dtype, A, ld, ns, B, ld, ns, h->Wm, M_out, A == B);
}

template <class T>
static int rotate_impl(cyx_ham_s* h, const void* A, const void* Q, void* out, int ns, int64_t ld) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  PhaseTimer tm(h->ctx, PH_ROTATE);
  hipLaunchKernelGGL(rotate_mfma_kernel<T>, dim3(ceil_div(N, (int64_t)RW * 16)), dim3(RW * 64), 0, h->ctx->stream,
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

template <class T>
static int resid_impl(cyx_ham_s* h, const void* A, const void* HA, const double* lam, int ns, int64_t ld, double* res) {
  hipLaunchKernelGGL(resnorm_kernel<T>, dim3(ns), dim3(1024), 0, h->ctx->stream, (const T*)A, (const T*)HA, lam, h->W,
                     h->N(), ld, res);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int ritz_residuals(cyx_ham_s* h, int dtype, const void* A, const void* HA, const double* lam, int ns, int64_t ld,
                   double* res) {
  switch (dtype) {
    case CYX_F64:
      return resid_impl<double>(h, A, HA, lam, ns, ld, res);
    case CYX_F32:
      return resid_impl<float>(h, A, HA, lam, ns, ld, res);
    case CYX_C128:
      return resid_impl<cplx<double>>(h, A, HA, lam, ns, ld, res);
    case CYX_C64:
      return resid_impl<cplx<float>>(h, A, HA, lam, ns, ld, res);
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
