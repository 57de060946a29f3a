// Subspace products of the mixed-precision path on the FP32 matrix cores
// (src/eigSolver.m:99-119 for complex64 / real FP32 blocks):
//
// * gram32: M = A^H diag(W) B with v_mfma_f32_16x16x4_f32.  The products
//   are exact and each MFMA accumulates in FP32, but only over G32_FLUSH k
//   steps (128 points); every such chunk is then added into FP64 registers,
//   and the waves' and blocks' partials are summed in FP64.  The FP32
//   rounding therefore enters only at the level of one 128-point chunk, and
//   the overlap keeps the accuracy the Rayleigh-Ritz step needs.  A lane
//   reads its four consecutive points of a column as 16-byte loads (with
//   8-byte loads per point the kernel ran 1.5x slower), and the tiles of one
//   k chunk run on one XCD.
// * rotate32: out = A Q with Q rounded to FP32 and the 256-term sums in
//   FP32.  Its error grows with how close the block's columns are to linear
//   dependence, so the caller uses it only when the overlap's Cholesky
//   factor certifies the block well conditioned.
//
// v_mfma_f32_16x16x4_f32 lane layout (checked on the device): lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and receives D[4(l/16) + t][l%16]
// in element t of its accumulator.
#include <cstdlib>
#include <string>

#include "cyx_internal.hpp"

namespace cyx {

typedef float v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4f mfma32(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int G32_TILE = 32;   // output tile of a wave (2 x 2 MFMA tiles)
constexpr int G32_SP = 4;      // points per lane per k step (16 points per step)
constexpr int G32_FLUSH = 8;   // k steps per FP32 chunk before the FP64 flush

__device__ inline void g32_tile_of(int tp, bool herm, int nta, int& ta, int& tb) {
  if (herm) {
    tb = 0;
    while ((tb + 1) * (tb + 2) / 2 <= tp) ++tb;
    ta = tp - tb * (tb + 1) / 2;
  } else {
    ta = tp % nta;
    tb = tp / nta;
  }
}

template <class T>
struct G32Step {
  T a[2][G32_SP], b[2][G32_SP];
  float w[G32_SP];
};

// the G32_SP consecutive points of one column that a lane takes, as 16-byte loads
template <class T>
__device__ __forceinline__ void g32_vec(T (&dst)[G32_SP], const T* __restrict__ p) {
  const v4f* q = reinterpret_cast<const v4f*>(p);
  if constexpr (traits<T>::is_complex) {
    const v4f u = q[0], v = q[1];
    dst[0] = T{u[0], u[1]};
    dst[1] = T{u[2], u[3]};
    dst[2] = T{v[0], v[1]};
    dst[3] = T{v[2], v[3]};
  } else {
    const v4f u = q[0];
#pragma unroll
    for (int s = 0; s < G32_SP; ++s) dst[s] = u[s];
  }
}

// vec: the leading dimensions are multiples of G32_SP and the blocks 16-byte
// aligned, so a lane's G32_SP points of a column are one aligned run
template <class T>
__device__ __forceinline__ void g32_load(G32Step<T>& S, const T* __restrict__ A, int64_t lda, int na,
                                         const T* __restrict__ B, int64_t ldb, int nb, const double* __restrict__ W,
                                         int64_t n0, int64_t k1, const int (&ca)[2], const int (&cb)[2], int g,
                                         bool vec) {
  const int64_t nl = n0 + G32_SP * g;
  if (vec && nl + G32_SP <= k1) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    const v2d* wq = reinterpret_cast<const v2d*>(W + nl);
    const v2d w0 = wq[0], w1 = wq[1];
    S.w[0] = (float)w0[0];
    S.w[1] = (float)w0[1];
    S.w[2] = (float)w1[0];
    S.w[3] = (float)w1[1];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (ca[i] < na) {
        g32_vec(S.a[i], A + (int64_t)ca[i] * lda + nl);
      } else {
#pragma unroll
        for (int s = 0; s < G32_SP; ++s) S.a[i][s] = zero<T>();
      }
      if (cb[i] < nb) {
        g32_vec(S.b[i], B + (int64_t)cb[i] * ldb + nl);
      } else {
#pragma unroll
        for (int s = 0; s < G32_SP; ++s) S.b[i][s] = zero<T>();
      }
    }
    return;
  }
#pragma unroll
  for (int s = 0; s < G32_SP; ++s) {
    const int64_t n = n0 + G32_SP * g + s;
    const bool ok = n < k1;
    S.w[s] = ok ? (float)W[n] : 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      S.a[i][s] = (ok && ca[i] < na) ? A[(int64_t)ca[i] * lda + n] : zero<T>();
      S.b[i][s] = (ok && cb[i] < nb) ? B[(int64_t)cb[i] * ldb + n] : zero<T>();
    }
  }
}

template <class T, int NC>
__device__ __forceinline__ void g32_mfma(const G32Step<T>& S, v4f (&acc)[NC][2][2]) {
#pragma unroll
  for (int s = 0; s < G32_SP; ++s) {
    if constexpr (traits<T>::is_complex) {
      // conj(a) w b: re += ar br + ai bi, im += ar bi - ai br; grouped by
      // product so that two MFMAs on one accumulator are 8 apart
      float br[2], bi[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        br[j] = S.w[s] * S.b[j][s].x;
        bi[j] = S.w[s] * S.b[j][s].y;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[0][i][j] = mfma32(S.a[i][s].x, br[j], acc[0][i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[1][i][j] = mfma32(S.a[i][s].x, bi[j], acc[1][i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[0][i][j] = mfma32(S.a[i][s].y, bi[j], acc[0][i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[1][i][j] = mfma32(-S.a[i][s].y, br[j], acc[1][i][j]);
    } else {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[0][i][j] = mfma32(S.a[i][s], S.w[s] * S.b[j][s], acc[0][i][j]);
    }
  }
}

template <int NC>
__device__ __forceinline__ void g32_flush(v4f (&acc)[NC][2][2], double (&accd)[NC][2][2][4]) {
#pragma unroll
  for (int q = 0; q < NC; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int t = 0; t < 4; ++t) accd[q][i][j][t] += (double)acc[q][i][j][t];
        acc[q][i][j] = v4f{0, 0, 0, 0};
      }
}

// One block of 4 waves: wave wv takes k sub-range (blockIdx.y * 4 + wv) of
// tile blockIdx.x; the four FP64 partials are summed in LDS in a fixed
// order (wave 3, 2, 1, 0) and written as one partial per block.
template <class T, bool FLUSH>
__global__ void __launch_bounds__(256) gram32_kernel(const T* __restrict__ A, int64_t lda, int na,
                                                     const T* __restrict__ B, int64_t ldb, int nb,
                                                     const double* __restrict__ W, int64_t N, int64_t chunk, int herm,
                                                     int nta, int ntp, int nchb, int vec,
                                                     typename acc_of<T>::type* __restrict__ part) {
  using D = typename acc_of<T>::type;
  constexpr bool CX = traits<T>::is_complex;
  constexpr int NC = CX ? 2 : 1;
  __shared__ double red[NC][2][2][4][64];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  // XCD-aware order: block b runs on XCD b % 8, so the tiles of chunk k are
  // given blocks congruent to k mod 8 and all run on one XCD, whose L2 then
  // serves the chunk's rows of A and B to every tile (nchb is a multiple of
  // 8; chunks past the end add zero partials)
  const int b = blockIdx.x;
  const int slot = b >> 3, tile = slot % ntp, kch = 8 * (slot / ntp) + (b & 7);
  int ta, tb;
  g32_tile_of(tile, herm != 0, nta, ta, tb);
  const int a0 = ta * G32_TILE, b0 = tb * G32_TILE;
  const int64_t k0 = ((int64_t)kch * 4 + wv) * chunk, k1 = min(N, k0 + chunk);
  v4f acc[NC][2][2];
  double accd[NC][2][2][4];
#pragma unroll
  for (int q = 0; q < NC; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[q][i][j] = v4f{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < 4; ++t) accd[q][i][j][t] = 0.0;
      }
  const int ca[2] = {a0 + c, a0 + 16 + c}, cb[2] = {b0 + c, b0 + 16 + c};
  constexpr int KSTEP = 4 * G32_SP;
  if (k0 < k1) {
    G32Step<T> S0, S1;
    g32_load(S0, A, lda, na, B, ldb, nb, W, k0, k1, ca, cb, g, vec != 0);
    int nstep = 0;
    (void)nstep;
    for (int64_t n0 = k0;;) {
      const bool more = n0 + KSTEP < k1;
      if (more) g32_load(S1, A, lda, na, B, ldb, nb, W, n0 + KSTEP, k1, ca, cb, g, vec != 0);
      __builtin_amdgcn_sched_barrier(0);
      g32_mfma(S0, acc);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (FLUSH) {
        if (++nstep == G32_FLUSH) {
          g32_flush(acc, accd);
          nstep = 0;
        }
      }
      if (!more) break;
      n0 += KSTEP;
      const bool more2 = n0 + KSTEP < k1;
      if (more2) g32_load(S0, A, lda, na, B, ldb, nb, W, n0 + KSTEP, k1, ca, cb, g, vec != 0);
      __builtin_amdgcn_sched_barrier(0);
      g32_mfma(S1, acc);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (FLUSH) {
        if (++nstep == G32_FLUSH) {
          g32_flush(acc, accd);
          nstep = 0;
        }
      }
      if (!more2) break;
      n0 += KSTEP;
    }
  }
  if constexpr (FLUSH) g32_flush(acc, accd);
  for (int w = 3; w >= 0; --w) {
    if (wv == w) {
#pragma unroll
      for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
              const double v = FLUSH ? accd[q][i][j][t] : (double)acc[q][i][j][t];
              red[q][i][j][t][l] = (w == 3) ? v : red[q][i][j][t][l] + v;
            }
    }
    __syncthreads();
  }
  D* out = part + ((int64_t)tile * nchb + kch) * (G32_TILE * G32_TILE);
  for (int e = threadIdx.x; e < G32_TILE * G32_TILE; e += 256) {
    const int ll = e & 63, t = (e >> 6) & 3, j = (e >> 8) & 1, i = (e >> 9) & 1;
    const int al = 16 * i + 4 * (ll >> 4) + t, bl = 16 * j + (ll & 15);
    D v{};
    if constexpr (CX)
      v = D{red[0][i][j][t][ll], red[1][i][j][t][ll]};
    else
      v = red[0][i][j][t][ll];
    out[bl * G32_TILE + al] = v;
  }
}

// M (na x nb column-major) = fixed-order FP64 sum of the block partials;
// Hermitian products keep the upper triangle and mirror it
template <class D>
__global__ void gram32_reduce_kernel(const D* __restrict__ part, int nchb, int herm, int nta, int na, int nb,
                                     D* __restrict__ M) {
  const int tp = blockIdx.x, e = threadIdx.x + blockDim.x * blockIdx.y;
  if (e >= G32_TILE * G32_TILE) return;
  int ta, tb;
  g32_tile_of(tp, herm != 0, nta, ta, tb);
  const int a = ta * G32_TILE + (e % G32_TILE), b = tb * G32_TILE + (e / G32_TILE);
  D s{};
  for (int k = 0; k < nchb; ++k) s = add(s, part[((int64_t)tp * nchb + k) * (G32_TILE * G32_TILE) + e]);
  if (a < na && b < nb) {
    if (!herm) {
      M[(int64_t)b * na + a] = s;
    } else if (a <= b) {
      M[(int64_t)b * na + a] = s;
      if (a != b) M[(int64_t)a * na + b] = conjg(s);
    }
  }
}

// ---- rotation out = A Q as out^T = Q^T A^T: lane (g, c) holds
// out[n0 + c][b0 + 4g + t], 16 consecutive points of one column per store.
// A block of R32W waves covers 16 R32W points and all ns columns in passes of
// R32B; Q[a chunk][R32B columns] is staged in LDS rounded to FP32 (passes of
// 128 columns, which read A half as often, measured slower: 2.05 against
// 1.60 ms at 276,480 x 256).
constexpr int R32W = 8;
constexpr int R32AC = 32;
constexpr int R32B = 64;
constexpr int R32BT = R32B / 16;

template <class T>
__global__ void __launch_bounds__(R32W * 64) rotate32_kernel(const T* __restrict__ A,
                                                             const typename acc_of<T>::type* __restrict__ Q,
                                                             T* __restrict__ out, int64_t N, int64_t ld, int ns) {
  constexpr bool CX = traits<T>::is_complex;
  constexpr int NC = CX ? 2 : 1;
  __shared__ T sq[R32AC][R32B + 1];  // [a][b]
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  const int64_t n = ((int64_t)blockIdx.x * R32W + wv) * 16 + c;
  const bool nok = n < N;
  for (int bg = 0; bg < ns; bg += R32B) {
    v4f acc[NC][R32BT];
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int bt = 0; bt < R32BT; ++bt) acc[q][bt] = v4f{0, 0, 0, 0};
    for (int ac = 0; ac < ns; ac += R32AC) {
      __syncthreads();
      for (int e = threadIdx.x; e < R32AC * R32B; e += R32W * 64) {
        const int aa = e % R32AC, bb = e / R32AC;
        sq[aa][bb] = (ac + aa < ns && bg + bb < ns) ? narrow<T>(Q[(int64_t)(bg + bb) * ns + ac + aa]) : zero<T>();
      }
      __syncthreads();
#pragma unroll
      for (int k4 = 0; k4 < R32AC; k4 += 4) {
        const int a = ac + k4 + g;
        const T x = (nok && a < ns) ? A[(int64_t)a * ld + n] : zero<T>();
        // (grouping the MFMAs by product, so that one accumulator's two
        // updates are further apart, measured 1.71 against 1.60 ms)
#pragma unroll
        for (int bt = 0; bt < R32BT; ++bt) {
          const T qv = sq[k4 + g][bt * 16 + c];
          if constexpr (CX) {
            // x q: re += xr qr - xi qi, im += xi qr + xr qi
            acc[0][bt] = mfma32(qv.x, x.x, acc[0][bt]);
            acc[1][bt] = mfma32(qv.x, x.y, acc[1][bt]);
            acc[0][bt] = mfma32(-qv.y, x.y, acc[0][bt]);
            acc[1][bt] = mfma32(qv.y, x.x, acc[1][bt]);
          } else {
            acc[0][bt] = mfma32(qv, x, acc[0][bt]);
          }
        }
      }
    }
    if (nok) {
#pragma unroll
      for (int bt = 0; bt < R32BT; ++bt)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int b = bg + 16 * bt + 4 * g + t;
          if (b < ns) {
            T v;
            if constexpr (CX)
              v = T{acc[0][bt][t], acc[1][bt][t]};
            else
              v = acc[0][bt][t];
            out[(int64_t)b * ld + n] = v;
          }
        }
    }
  }
}

// ---------------------------------------------------------------- host side
template <class T>
static int gram32_impl(cyx_ham_s* h, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                       const double* W, void* M_out, bool herm) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  herm = herm && na == nb;
  const int nta = ceil_div(na, G32_TILE), ntb = ceil_div(nb, G32_TILE);
  const int ntp = herm ? nta * (nta + 1) / 2 : nta * ntb;
  // about 8192 waves in all (16384 for a Hermitian product, whose upper
  // tiles alone leave the machine short of waves), k sub-ranges in whole steps
  // (FP32 sums over whole 512-point wave ranges without the FP64 flush
  // measured the same time, so the flush stays)
  int nchb = std::max(1, ceil_div(herm ? 16384 : 8192, 4 * ntp));
  const int kstep = 4 * G32_SP;
  int64_t chunk = ceil_div(N, 4 * (int64_t)nchb);
  chunk = ceil_div(chunk, kstep) * kstep;
  nchb = (int)ceil_div(N, 4 * chunk);
  nchb = ceil_div(nchb, 8) * 8;  // the XCD-aware block order takes chunks eight at a time
  const size_t pbytes = (size_t)ntp * nchb * G32_TILE * G32_TILE * sizeof(D);
  CYX_CHECK(ensure(h->small, pbytes) == CYX_OK, "partials allocation failed");
  hipStream_t st = h->ctx->stream;
  PhaseTimer tm(h->ctx, PH_GRAM);
  const int nblk = ntp * nchb;
  // 16-byte vector loads when every lane's run of G32_SP points is aligned
  const auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const int vec = (lda % G32_SP == 0 && ldb % G32_SP == 0 && al(A) && al(B) && ((uintptr_t)W & 31) == 0) ? 1 : 0;
  hipLaunchKernelGGL((gram32_kernel<T, true>), dim3(nblk), dim3(256), 0, st, (const T*)A, lda, na, (const T*)B, ldb,
                     nb, W, N, chunk, herm ? 1 : 0, nta, ntp, nchb, vec, (D*)h->small.p);
  CYX_HIP(hipGetLastError());
  hipLaunchKernelGGL(gram32_reduce_kernel<D>, dim3(ntp, G32_TILE * G32_TILE / 256), dim3(256), 0, st,
                     (const D*)h->small.p, nchb, herm ? 1 : 0, nta, na, nb, (D*)M_out);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int gram_rect32(cyx_ham_s* h, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                const double* W, void* M_out, bool herm) {
  switch (dtype) {
    case CYX_F32:
      return gram32_impl<float>(h, A, lda, na, B, ldb, nb, W, M_out, herm);
    case CYX_C64:
      return gram32_impl<cplx<float>>(h, A, lda, na, B, ldb, nb, W, M_out, herm);
  }
  set_error("gram_rect32 takes FP32 blocks only (dtype %d)", dtype);
  return CYX_ERR_ARG;
}

template <class T>
static int rotate32_impl(cyx_ham_s* h, const void* A, const void* Q, void* out, int ns, int64_t ld) {
  using D = typename acc_of<T>::type;
  const int64_t N = h->N();
  PhaseTimer tm(h->ctx, PH_ROTATE);
  hipLaunchKernelGGL(rotate32_kernel<T>, dim3(ceil_div(N, (int64_t)R32W * 16)), dim3(R32W * 64), 0, h->ctx->stream,
                     (const T*)A, (const D*)Q, (T*)out, N, ld, ns);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int rotate32(cyx_ham_s* h, int dtype, const void* A, const void* Q, void* out, int ns, int64_t ld) {
  switch (dtype) {
    case CYX_F32:
      return rotate32_impl<float>(h, A, Q, out, ns, ld);
    case CYX_C64:
      return rotate32_impl<cplx<float>>(h, A, Q, out, ns, ld);
  }
  set_error("rotate32 takes FP32 blocks only (dtype %d)", dtype);
  return CYX_ERR_ARG;
}

// min over i of R_ii^2 / M_ii for the Cholesky factor R (upper, in place of
// M) and the saved diagonal of M: 1 for orthogonal columns of any scale,
// small when a column lies close to the span of the earlier ones
template <class D>
__global__ void chol_ratio_kernel(const D* __restrict__ R, const double* __restrict__ mdiag, int ns,
                                  double* __restrict__ out) {
  __shared__ double red[256];
  double m = 1.0;
  for (int i = threadIdx.x; i < ns; i += 256) {
    double r;
    if constexpr (traits<D>::is_complex)
      r = R[(int64_t)i * ns + i].x;
    else
      r = R[(int64_t)i * ns + i];
    const double d = mdiag[i];
    m = fmin(m, d > 0 ? r * r / d : 0.0);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <class D>
__global__ void diag_copy_kernel(const D* __restrict__ M, int ns, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  if constexpr (traits<D>::is_complex)
    out[i] = M[(int64_t)i * ns + i].x;
  else
    out[i] = M[(int64_t)i * ns + i];
}

int overlap_diag(cyx_ham_s* h, bool cx, const void* M, int ns, double* out) {
  if (cx)
    hipLaunchKernelGGL(diag_copy_kernel<cplx<double>>, dim3(ceil_div(ns, 256)), dim3(256), 0, h->ctx->stream,
                       (const cplx<double>*)M, ns, out);
  else
    hipLaunchKernelGGL(diag_copy_kernel<double>, dim3(ceil_div(ns, 256)), dim3(256), 0, h->ctx->stream,
                       (const double*)M, ns, out);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int cholesky_ratio(cyx_ham_s* h, bool cx, const void* R, const double* mdiag, int ns, double* out) {
  if (cx)
    hipLaunchKernelGGL(chol_ratio_kernel<cplx<double>>, dim3(1), dim3(256), 0, h->ctx->stream,
                       (const cplx<double>*)R, mdiag, ns, out);
  else
    hipLaunchKernelGGL(chol_ratio_kernel<double>, dim3(1), dim3(256), 0, h->ctx->stream, (const double*)R, mdiag,
                       ns, out);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

}  // namespace cyx
