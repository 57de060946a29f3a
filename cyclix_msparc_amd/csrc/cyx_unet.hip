// Learned part of the Poisson preconditioner: a linear three-level 3-D
// convolutional U-Net on the cylindrical grid (r, theta, z), BF16 weights
// and BF16 matrix-core products with FP32 accumulation and FP32 activations.
//
//   level 0 (nr, nt, nz), level 1 (nr/2, nt/2, nz/2), level 2 (nr/4, nt/4, nz/4)
//   a0 = conv3(x, 1->16)          L0      c1 = up(c2, 64->32) + b1    L1
//   a1 = down(a0, 16->32)         L1      g1 = conv3(c1, 32->32)      L1
//   b1 = conv3(a1, 32->32)        L1      c0 = up(g1, 32->16) + a0    L0
//   a2 = down(b1, 32->64)         L2      y  = conv3(c0, 16->1)       L0
//   b2 = conv3(a2, 64->64)        L2      U(x) = s[r] * y
//   c2 = conv3(b2, 64->64)        L2
//
// conv3: 3x3x3, stride 1, zero padding in r (the Dirichlet end), periodic in
// theta and z; down: 2x2x2 stride 2; up: its transpose.  No activations, so
// U is a linear operator.  AAR uses d = D^{-1} r + U(D^{-1} r) (D the Jacobi
// diagonal of Lap_std) as a correction direction with exact line search after
// its Anderson iterations (cyx_aar.hip).
//
// Training: per-sample normalised squared error of U against the target,
// backward through the same kernels (the adjoint of conv3 is conv3 with the
// flipped, transposed kernel; of down is up and of up is down with the
// transposed kernel), FP32 weight gradients and Adam on FP32 master weights.
#include "cyx_internal.hpp"
#include "cyx_reduce.hpp"

namespace cyx {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

enum ConvMode { M_CONV3 = 0, M_DOWN = 1, M_UP = 2 };

struct Lvl {
  int nr, nt, nz;
};
__host__ __device__ inline int64_t npts(const Lvl& L) { return (int64_t)L.nr * L.nt * L.nz; }

__device__ __forceinline__ int wrapn(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// conv3 tap t = (dz+1)*9 + (dt+1)*3 + (dr+1); down/up tap t = c*4 + b*2 + a
// for the fine offsets (z, theta, r) = (c, b, a)

// ---------------------------------------------------------------- FMA convolution (any mode, small layers)
// out[b][q][co] (=|+=) sum W[tap][co][ci] in[b][p(q, tap)][ci] (+ skip[b][q][co])
// One thread per (output point, group of COB output channels); the weights
// are staged in LDS as [tap][ci][co] so that a thread's COB weights of one
// (tap, ci) are contiguous (wide broadcast reads).
template <int MODE, int COB>
__global__ void __launch_bounds__(256) conv_fma_kernel(const float* __restrict__ in, int cin, Lvl li,
                                                       float* __restrict__ out, int cout, Lvl lo,
                                                       const unsigned short* __restrict__ W,
                                                       const float* __restrict__ skip, int accumulate, int nb) {
  extern __shared__ __align__(16) unsigned short wsh[];
  constexpr int NT = MODE == M_CONV3 ? 27 : 8;
  // per-tap stride padded by 16 elements (32 B): in up mode the lanes of a
  // wave read up to 8 taps at once, which then fall in different banks
  const int tstride = cin * cout + 16;
  for (int tap = 0; tap < NT; ++tap)
    for (int t = threadIdx.x; t < cin * cout; t += blockDim.x) {
      const int co = t / cin, ci = t - co * cin;
      wsh[tap * tstride + ci * cout + co] = W[tap * cin * cout + t];
    }
  __syncthreads();
  const int ncb = cout / COB;
  const int64_t no = npts(lo), ni = npts(li);
  const int64_t gq = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gq >= no * nb * ncb) return;
  const int cb = (int)(gq % ncb);
  const int64_t rest = gq / ncb;
  const int b = (int)(rest / no);
  const int q = (int)(rest - (int64_t)b * no);
  const int i = q % lo.nr, j = (q / lo.nr) % lo.nt, k = q / (lo.nr * lo.nt);
  float acc[COB];
#pragma unroll
  for (int c = 0; c < COB; ++c) acc[c] = 0.f;
  const float* ib = in + (int64_t)b * ni * cin;
  const unsigned short* wc = wsh + cb * COB;
  if constexpr (MODE == M_UP) {
    const int tap = (k & 1) * 4 + (j & 1) * 2 + (i & 1);
    const int p = ((k >> 1) * li.nt + (j >> 1)) * li.nr + (i >> 1);
    for (int ci = 0; ci < cin; ++ci) {
      const float v = ib[(int64_t)p * cin + ci];
      const unsigned short* w = wc + tap * tstride + ci * cout;
#pragma unroll
      for (int c = 0; c < COB; ++c) acc[c] = __builtin_fmaf(bf2f(w[c]), v, acc[c]);
    }
  } else {
    for (int tap = 0; tap < NT; ++tap) {
      int ii, jj, kk;
      if constexpr (MODE == M_CONV3) {
        ii = i + tap % 3 - 1;
        jj = wrapn(j + (tap / 3) % 3 - 1, li.nt);
        kk = wrapn(k + tap / 9 - 1, li.nz);
        if (ii < 0 || ii >= li.nr) continue;
      } else {
        ii = 2 * i + (tap & 1);
        jj = 2 * j + ((tap >> 1) & 1);
        kk = 2 * k + (tap >> 2);
      }
      const int p = (kk * li.nt + jj) * li.nr + ii;
      for (int ci = 0; ci < cin; ++ci) {
        const float v = ib[(int64_t)p * cin + ci];
        const unsigned short* w = wc + tap * tstride + ci * cout;
#pragma unroll
        for (int c = 0; c < COB; ++c) acc[c] = __builtin_fmaf(bf2f(w[c]), v, acc[c]);
      }
    }
  }
  const int64_t ob = ((int64_t)b * no + q) * cout + cb * COB;
#pragma unroll
  for (int c = 0; c < COB; ++c) {
    float v = acc[c];
    if (skip) v += skip[ob + c];
    if (accumulate) v += out[ob + c];
    out[ob + c] = v;
  }
}

// ---------------------------------------------------------------- MFMA conv3 (16 | 32 | 64 channels)
// One wave computes 16 consecutive radial points x COUT channels with
// v_mfma_f32_16x16x16_bf16: A (16 points x 16 input channels of one tap),
// B (16 input channels x 16 output channels of that tap).  Lane l holds
// A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16] and D[4(l/16)+t][l%16].
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) conv3_mfma_kernel(const float* __restrict__ in, float* __restrict__ out, Lvl L,
                                                         const unsigned short* __restrict__ W,
                                                         const float* __restrict__ skip, int accumulate) {
  // block: 4 waves on 4 consecutive theta rows; grid.x: (r tile, output-channel tile)
  constexpr int NTL = COUT / 16;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int t = blockIdx.x % NTL;
  const int i0 = (blockIdx.x / NTL) * 16;
  const int j = blockIdx.y * 4 + wave;
  const int k = blockIdx.z % L.nz;
  const int b = blockIdx.z / L.nz;
  if (j >= L.nt) return;
  const int m = lane & 15, kg = lane >> 4;
  const int64_t np = npts(L);
  const float* ib = in + (int64_t)b * np * CIN;
  v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
  const bool row_ok = i0 + m < L.nr;
#pragma unroll 3
  for (int tap = 0; tap < 27; ++tap) {
    const int ii = i0 + m + tap % 3 - 1;
    const int jj = wrapn(j + (tap / 3) % 3 - 1, L.nt);
    const int kk = wrapn(k + tap / 9 - 1, L.nz);
    const bool ok = row_ok && ii >= 0 && ii < L.nr;
    const int64_t p = ((int64_t)kk * L.nt + jj) * L.nr + (ok ? ii : 0);
    const unsigned short* wt = W + ((int64_t)(tap * COUT + t * 16 + m) * CIN + kg * 4);
#pragma unroll
    for (int kb = 0; kb < CIN / 16; ++kb) {
      const int ci0 = kb * 16 + kg * 4;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (ok) v = *reinterpret_cast<const float4*>(ib + p * CIN + ci0);
      v4s a;
      a[0] = (short)f2bf(v.x);
      a[1] = (short)f2bf(v.y);
      a[2] = (short)f2bf(v.z);
      a[3] = (short)f2bf(v.w);
      const uint2 wv = *reinterpret_cast<const uint2*>(wt + kb * 16);
      v4s bb;
      bb[0] = (short)(wv.x & 0xffff);
      bb[1] = (short)(wv.x >> 16);
      bb[2] = (short)(wv.y & 0xffff);
      bb[3] = (short)(wv.y >> 16);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int i = i0 + kg * 4 + e;
    if (i < L.nr) {
      const int64_t q = ((int64_t)k * L.nt + j) * L.nr + i;
      const int64_t o = ((int64_t)b * np + q) * COUT + t * 16 + m;
      float v = acc[e];
      if (skip) v += skip[o];
      if (accumulate) v += out[o];
      out[o] = v;
    }
  }
}

// ---------------------------------------------------------------- weight gradients
// dW[tap][co][ci] += sum_b sum_{(q,p) of tap} dout[b][q][co] in[b][p][ci]
// Pairs are enumerated over the "driver" level: the output level for conv3
// and down, the input (coarse) level for up.  Each block: one tap, a chunk of
// driver points, all batch samples; per-thread register accumulators over a
// strided set of (co, ci) pairs, staged 16 points at a time through LDS.
template <int MODE>
__global__ void __launch_bounds__(256) conv_dw_kernel(const float* __restrict__ dout, int cout, Lvl lo,
                                                      const float* __restrict__ in, int cin, Lvl li,
                                                      float* __restrict__ dW, int nb, int chunk) {
  constexpr int PT = 16;
  constexpr int MAXP = 16;  // pairs per thread: cout*cin <= 4096
  __shared__ float sd[PT][64];
  __shared__ float si[PT][64];
  const int tap = blockIdx.y;
  const Lvl ld = MODE == M_UP ? li : lo;
  const int64_t ndrv = npts(ld);
  const int64_t q0 = (int64_t)blockIdx.x * chunk;
  const int64_t q1 = min(ndrv, q0 + chunk);
  const int npair = cout * cin;
  float acc[MAXP];
#pragma unroll
  for (int e = 0; e < MAXP; ++e) acc[e] = 0.f;
  const int64_t no = npts(lo), ni = npts(li);
  for (int b = 0; b < nb; ++b) {
    for (int64_t qb = q0; qb < q1; qb += PT) {
      // stage PT pairs: dout row (cout) and in row (cin), zero when invalid
      for (int t = threadIdx.x; t < PT * 64; t += 256) {
        const int pt = t / 64, c = t % 64;
        const int64_t qd = qb + pt;
        float vd = 0.f, vi = 0.f;
        if (qd < q1) {
          const int d = (int)qd;
          const int i = d % ld.nr, j = (d / ld.nr) % ld.nt, k = d / (ld.nr * ld.nt);
          int64_t qo = -1, pi = -1;
          if constexpr (MODE == M_CONV3) {
            const int ii = i + tap % 3 - 1;
            if (ii >= 0 && ii < li.nr) {
              const int jj = wrapn(j + (tap / 3) % 3 - 1, li.nt), kk = wrapn(k + tap / 9 - 1, li.nz);
              qo = d;
              pi = ((int64_t)kk * li.nt + jj) * li.nr + ii;
            }
          } else if constexpr (MODE == M_DOWN) {
            qo = d;
            pi = ((int64_t)(2 * k + (tap >> 2)) * li.nt + 2 * j + ((tap >> 1) & 1)) * li.nr + 2 * i + (tap & 1);
          } else {
            pi = d;
            qo = ((int64_t)(2 * k + (tap >> 2)) * lo.nt + 2 * j + ((tap >> 1) & 1)) * lo.nr + 2 * i + (tap & 1);
          }
          if (qo >= 0) {
            if (c < cout) vd = dout[((int64_t)b * no + qo) * cout + c];
            if (c < cin) vi = in[((int64_t)b * ni + pi) * cin + c];
          }
        }
        sd[pt][c] = vd;
        si[pt][c] = vi;
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < MAXP; ++e) {
        const int pr = threadIdx.x + 256 * e;
        if (pr < npair) {
          const int co = pr / cin, ci = pr % cin;
          float s = acc[e];
#pragma unroll
          for (int pt = 0; pt < PT; ++pt) s = __builtin_fmaf(sd[pt][co], si[pt][ci], s);
          acc[e] = s;
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int e = 0; e < MAXP; ++e) {
    const int pr = threadIdx.x + 256 * e;
    if (pr < npair) atomicAdd(&dW[(int64_t)tap * npair + pr], acc[e]);
  }
}

// ---------------------------------------------------------------- small kernels
// Wt[map(tap)][ci][co] = W[tap][co][ci]; map = 26 - tap when flip (conv3 adjoint)
__global__ void transpose_w_kernel(const unsigned short* __restrict__ W, unsigned short* __restrict__ Wt, int ntap,
                                   int cout, int cin, int flip) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int tot = ntap * cout * cin;
  if (e >= tot) return;
  const int tap = e / (cout * cin);
  const int co = (e / cin) % cout;
  const int ci = e % cin;
  const int t2 = flip ? ntap - 1 - tap : tap;
  Wt[((int64_t)t2 * cin + ci) * cout + co] = W[e];
}

__global__ void cast_bf16_kernel(const float* __restrict__ p, unsigned short* __restrict__ q, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) q[e] = f2bf(p[e]);
}

// x[b][n] = float(src[b][n] * scale_b / diag[r])  (input of U: the Jacobi-scaled residual)
template <class S>
__global__ void unet_input_kernel(const S* __restrict__ src, const double* __restrict__ coef, int nr, int64_t N,
                                  int nb, const float* __restrict__ scale, float* __restrict__ x) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * nb) return;
  const int64_t n = e % N;
  const int b = (int)(e / N);
  const double d = coef[(n % nr) * RCOEF + RCOEF - 1];
  const double s = scale ? (double)scale[b] : 1.0;
  x[e] = (float)((double)src[e] * s / d);
}

// per-sample sum of squares (b = blockIdx.y), partial per block
__global__ void __launch_bounds__(AB) sumsq_kernel(const float* __restrict__ a, int64_t N, double* part) {
  __shared__ double red[AB];
  const int b = blockIdx.y;
  double s = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const double v = a[(int64_t)b * N + n];
    s += v * v;
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) part[(int64_t)b * gridDim.x + blockIdx.x] = t;
}

// u = s[r] y;  e = u - t;  dy = s[r] * 2 e * w_b;  ds[r] += 2 e y w_b;  loss partial sum e^2 w_b
// (w_b = 1 / (nb * ||t_b||^2): mean over samples of the relative squared error)
__global__ void __launch_bounds__(AB) head_loss_kernel(const float* __restrict__ y, const float* __restrict__ sr,
                                                       const float* __restrict__ tgt, const double* __restrict__ wb,
                                                       int nr, int64_t N, float* __restrict__ dy,
                                                       float* __restrict__ ds, double* part) {
  __shared__ double red[AB];
  __shared__ float dsl[512];
  for (int t = threadIdx.x; t < nr; t += AB) dsl[t] = 0.f;
  __syncthreads();
  const int b = blockIdx.y;
  const double w = wb[b];
  double l = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const int r = (int)(n % nr);
    const int64_t o = (int64_t)b * N + n;
    const float yy = y[o];
    const double e = (double)sr[r] * yy - (double)tgt[o];
    l += e * e * w;
    dy[o] = (float)(2.0 * e * w * sr[r]);
    atomicAdd(&dsl[r], (float)(2.0 * e * w * yy));
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nr; t += AB)
    if (dsl[t] != 0.f) atomicAdd(&ds[t], dsl[t]);
  const double s = block_sum(l, red);
  if (threadIdx.x == 0) part[(int64_t)b * gridDim.x + blockIdx.x] = s;
}

// out[b][n] = s[r] y[b][n]
__global__ void head_apply_kernel(const float* __restrict__ y, const float* __restrict__ sr, int nr, int64_t N,
                                  int nb, float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < N * nb) out[e] = sr[(e % N) % nr] * y[e];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float a = b1 * m1[e] + (1.f - b1) * gg;
  const float v = b2 * m2[e] + (1.f - b2) * gg * gg;
  m1[e] = a;
  m2[e] = v;
  p[e] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ void zero_f_kernel(float* p, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0.f;
}

}  // namespace cyx

using namespace cyx;

// ---------------------------------------------------------------- network object
namespace {
struct LayerSpec {
  int mode, cin, cout, lvl;  // lvl: level of the layer's output (conv3, up) or input (down)
};
// e0 d1 e1 d2 e2 m2 u1 f1 u0 f0
constexpr int NLAYER = 10;
constexpr LayerSpec SPEC[NLAYER] = {
    {M_CONV3, 1, 16, 0},  {M_DOWN, 16, 32, 0},  {M_CONV3, 32, 32, 1}, {M_DOWN, 32, 64, 1}, {M_CONV3, 64, 64, 2},
    {M_CONV3, 64, 64, 2}, {M_UP, 64, 32, 1},    {M_CONV3, 32, 32, 1}, {M_UP, 32, 16, 0},   {M_CONV3, 16, 1, 0},
};
inline int ntap(int mode) { return mode == M_CONV3 ? 27 : 8; }
}  // namespace

struct cyx_unet_s {
  cyx_ctx_s* ctx = nullptr;
  Lvl lv[3];
  int64_t woff[NLAYER + 1];  // weight offsets; the radial scale s[nr] follows at woff[NLAYER]
  int64_t nparam = 0;
  float* p = nullptr;    // FP32 master parameters
  float* g = nullptr;    // gradients
  float* m1 = nullptr;   // Adam moments
  float* m2 = nullptr;
  unsigned short* wb = nullptr;   // BF16 weights
  unsigned short* wt = nullptr;   // transposed BF16 weights (backward data)
  int64_t adam_t = 0;
  int cap = 0;           // batch capacity of the activation buffers
  float* act = nullptr;  // activations (forward) and their gradients (backward)
  size_t act_bytes = 0;
  double* red = nullptr; // reduction partials / per-sample weights
  int64_t N() const { return npts(lv[0]); }
};

namespace {
// activation buffers of a batch: forward x a0 a1 b1 a2 b2 c2 c1 g1 c0 y,
// gradients dx(unused) da0 da1 db1 da2 db2 dc2 dc1 dg1 dc0 dy
enum { X = 0, A0, A1, B1, A2, B2, C2, C1, G1, C0, Y, NACT };
inline int act_ch(int a) {
  switch (a) {
    case X: return 1;
    case A0: return 16;
    case A1: case B1: case C1: case G1: return 32;
    case A2: case B2: case C2: return 64;
    case C0: return 16;
    default: return 1;
  }
}
inline int act_lvl(int a) {
  switch (a) {
    case A1: case B1: case C1: case G1: return 1;
    case A2: case B2: case C2: return 2;
    default: return 0;
  }
}

struct Acts {
  float* f[NACT];
  float* d[NACT];
};

Acts acts_of(cyx_unet_s* net, int nb) {
  Acts A{};
  float* base = net->act;
  for (int pass = 0; pass < 2; ++pass)
    for (int a = 0; a < NACT; ++a) {
      float*& slot = pass == 0 ? A.f[a] : A.d[a];
      slot = base;
      base += (size_t)nb * npts(net->lv[act_lvl(a)]) * act_ch(a);
    }
  return A;
}
size_t acts_bytes(cyx_unet_s* net, int nb) {
  size_t n = 0;
  for (int a = 0; a < NACT; ++a) n += (size_t)nb * npts(net->lv[act_lvl(a)]) * act_ch(a);
  return 2 * n * sizeof(float);
}

int ensure_acts(cyx_unet_s* net, int nb) {
  if (nb <= net->cap) return CYX_OK;
  const size_t need = acts_bytes(net, nb);
  if (net->act) {
    (void)hipStreamSynchronize(net->ctx->stream);
    (void)hipFree(net->act);
  }
  net->act = nullptr;
  net->cap = 0;
  if (hipMalloc((void**)&net->act, need) != hipSuccess) {
    set_error("U-Net activation allocation of %zu bytes failed", need);
    return CYX_ERR_NOMEM;
  }
  net->act_bytes = need;
  net->cap = nb;
  return CYX_OK;
}

template <int MODE>
int launch_fma(cyx_unet_s* net, const float* in, int cin, Lvl li, float* out, int cout, Lvl lo,
               const unsigned short* W, const float* skip, int accumulate, int nb) {
  hipStream_t st = net->ctx->stream;
  const size_t shm = (size_t)ntap(MODE) * (cin * cout + 16) * sizeof(unsigned short);
  if (shm > 64 * 1024) {
    set_error("internal: FMA convolution weights exceed LDS (%zu bytes)", shm);
    return CYX_ERR_ARG;
  }
  const int cob = cout >= 16 ? 16 : cout;
  if (cout % cob) {
    set_error("internal: unsupported channel count %d", cout);
    return CYX_ERR_ARG;
  }
  const int64_t tot = npts(lo) * nb * (cout / cob);
  const dim3 grid((unsigned)ceil_div(tot, 256));
  if (cob == 16)
    hipLaunchKernelGGL((conv_fma_kernel<MODE, 16>), grid, dim3(256), shm, st, in, cin, li, out, cout, lo, W, skip,
                       accumulate, nb);
  else if (cob == 1)
    hipLaunchKernelGGL((conv_fma_kernel<MODE, 1>), grid, dim3(256), shm, st, in, cin, li, out, cout, lo, W, skip,
                       accumulate, nb);
  else {
    set_error("internal: unsupported channel count %d", cout);
    return CYX_ERR_ARG;
  }
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int conv3(cyx_unet_s* net, const float* in, int cin, float* out, int cout, int lvl, const unsigned short* W,
          const float* skip, int accumulate, int nb) {
  const Lvl L = net->lv[lvl];
  hipStream_t st = net->ctx->stream;
  const dim3 grid((unsigned)(ceil_div(L.nr, 16) * (cout / 16)), (unsigned)ceil_div(L.nt, 4), (unsigned)(L.nz * nb));
  if (cin == 32 && cout == 32) {
    hipLaunchKernelGGL((conv3_mfma_kernel<32, 32>), grid, dim3(256), 0, st, in, out, L, W, skip, accumulate);
  } else if (cin == 64 && cout == 64) {
    hipLaunchKernelGGL((conv3_mfma_kernel<64, 64>), grid, dim3(256), 0, st, in, out, L, W, skip, accumulate);
  } else if (cin == 16 && cout == 16) {
    hipLaunchKernelGGL((conv3_mfma_kernel<16, 16>), grid, dim3(256), 0, st, in, out, L, W, skip, accumulate);
  } else {
    return launch_fma<M_CONV3>(net, in, cin, L, out, cout, L, W, skip, accumulate, nb);
  }
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int forward(cyx_unet_s* net, const Acts& A, int nb) {
  const unsigned short* w = net->wb;
  const int64_t* o = net->woff;
  const Lvl* lv = net->lv;
  int rc;
#define CK(x) \
  if ((rc = (x)) != CYX_OK) return rc
  CK(conv3(net, A.f[X], 1, A.f[A0], 16, 0, w + o[0], nullptr, 0, nb));
  CK(launch_fma<M_DOWN>(net, A.f[A0], 16, lv[0], A.f[A1], 32, lv[1], w + o[1], nullptr, 0, nb));
  CK(conv3(net, A.f[A1], 32, A.f[B1], 32, 1, w + o[2], nullptr, 0, nb));
  CK(launch_fma<M_DOWN>(net, A.f[B1], 32, lv[1], A.f[A2], 64, lv[2], w + o[3], nullptr, 0, nb));
  CK(conv3(net, A.f[A2], 64, A.f[B2], 64, 2, w + o[4], nullptr, 0, nb));
  CK(conv3(net, A.f[B2], 64, A.f[C2], 64, 2, w + o[5], nullptr, 0, nb));
  CK(launch_fma<M_UP>(net, A.f[C2], 64, lv[2], A.f[C1], 32, lv[1], w + o[6], A.f[B1], 0, nb));
  CK(conv3(net, A.f[C1], 32, A.f[G1], 32, 1, w + o[7], nullptr, 0, nb));
  CK(launch_fma<M_UP>(net, A.f[G1], 32, lv[1], A.f[C0], 16, lv[0], w + o[8], A.f[A0], 0, nb));
  CK(conv3(net, A.f[C0], 16, A.f[Y], 1, 0, w + o[9], nullptr, 0, nb));
#undef CK
  return CYX_OK;
}

// transposed (and for conv3 flipped) BF16 weights of layer l into net->wt
const unsigned short* wt_of(cyx_unet_s* net, int l) {
  const LayerSpec& s = SPEC[l];
  const int nt = ntap(s.mode);
  const int tot = nt * s.cin * s.cout;
  unsigned short* dst = net->wt + net->woff[l];
  hipLaunchKernelGGL(transpose_w_kernel, dim3(ceil_div(tot, 256)), dim3(256), 0, net->ctx->stream,
                     (const unsigned short*)(net->wb + net->woff[l]), dst, nt, s.cout, s.cin,
                     s.mode == M_CONV3 ? 1 : 0);
  return dst;
}

template <int MODE>
void launch_dw(cyx_unet_s* net, const float* dout, int cout, Lvl lo, const float* in, int cin, Lvl li, float* dW,
               int nb) {
  const Lvl ld = MODE == M_UP ? li : lo;
  const int64_t nd = npts(ld);
  const int chunk = 2048;
  const dim3 grid((unsigned)ceil_div(nd, chunk), (unsigned)ntap(MODE));
  hipLaunchKernelGGL(conv_dw_kernel<MODE>, grid, dim3(256), 0, net->ctx->stream, dout, cout, lo, in, cin, li, dW, nb,
                     chunk);
}

// gradients of all weights given dY (in A.d[Y]); activation gradients in A.d
int backward(cyx_unet_s* net, const Acts& A, int nb) {
  const Lvl* lv = net->lv;
  const int64_t* o = net->woff;
  float* g = net->g;
  int rc;
#define CK(x) \
  if ((rc = (x)) != CYX_OK) return rc
  // f0: y = conv3(c0)
  launch_dw<M_CONV3>(net, A.d[Y], 1, lv[0], A.f[C0], 16, lv[0], g + o[9], nb);
  CK(conv3(net, A.d[Y], 1, A.d[C0], 16, 0, wt_of(net, 9), nullptr, 0, nb));
  // u0: c0 = up(g1) + a0
  launch_dw<M_UP>(net, A.d[C0], 16, lv[0], A.f[G1], 32, lv[1], g + o[8], nb);
  CK(launch_fma<M_DOWN>(net, A.d[C0], 16, lv[0], A.d[G1], 32, lv[1], wt_of(net, 8), nullptr, 0, nb));
  // f1: g1 = conv3(c1)
  launch_dw<M_CONV3>(net, A.d[G1], 32, lv[1], A.f[C1], 32, lv[1], g + o[7], nb);
  CK(conv3(net, A.d[G1], 32, A.d[C1], 32, 1, wt_of(net, 7), nullptr, 0, nb));
  // u1: c1 = up(c2) + b1  (d b1 starts from d c1)
  launch_dw<M_UP>(net, A.d[C1], 32, lv[1], A.f[C2], 64, lv[2], g + o[6], nb);
  CK(launch_fma<M_DOWN>(net, A.d[C1], 32, lv[1], A.d[C2], 64, lv[2], wt_of(net, 6), nullptr, 0, nb));
  // m2: c2 = conv3(b2)
  launch_dw<M_CONV3>(net, A.d[C2], 64, lv[2], A.f[B2], 64, lv[2], g + o[5], nb);
  CK(conv3(net, A.d[C2], 64, A.d[B2], 64, 2, wt_of(net, 5), nullptr, 0, nb));
  // e2: b2 = conv3(a2)
  launch_dw<M_CONV3>(net, A.d[B2], 64, lv[2], A.f[A2], 64, lv[2], g + o[4], nb);
  CK(conv3(net, A.d[B2], 64, A.d[A2], 64, 2, wt_of(net, 4), nullptr, 0, nb));
  // d2: a2 = down(b1);  d b1 = d c1 + down^T(d a2)
  launch_dw<M_DOWN>(net, A.d[A2], 64, lv[2], A.f[B1], 32, lv[1], g + o[3], nb);
  CK(launch_fma<M_UP>(net, A.d[A2], 64, lv[2], A.d[B1], 32, lv[1], wt_of(net, 3), A.d[C1], 0, nb));
  // e1: b1 = conv3(a1)
  launch_dw<M_CONV3>(net, A.d[B1], 32, lv[1], A.f[A1], 32, lv[1], g + o[2], nb);
  CK(conv3(net, A.d[B1], 32, A.d[A1], 32, 1, wt_of(net, 2), nullptr, 0, nb));
  // d1: a1 = down(a0);  d a0 = d c0 + down^T(d a1)
  launch_dw<M_DOWN>(net, A.d[A1], 32, lv[1], A.f[A0], 16, lv[0], g + o[1], nb);
  CK(launch_fma<M_UP>(net, A.d[A1], 32, lv[1], A.d[A0], 16, lv[0], wt_of(net, 1), A.d[C0], 0, nb));
  // e0: a0 = conv3(x)
  launch_dw<M_CONV3>(net, A.d[A0], 16, lv[0], A.f[X], 1, lv[0], g + o[0], nb);
#undef CK
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int refresh_bf16(cyx_unet_s* net) {
  const int n = (int)net->woff[NLAYER];
  hipLaunchKernelGGL(cast_bf16_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, net->ctx->stream, (const float*)net->p,
                     net->wb, n);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

// loss and dY for targets t (nb x N); fills net->g's radial-scale slot; returns the loss on the host
int head_loss(cyx_unet_s* net, const Acts& A, const float* tgt, int nb, double* loss) {
  hipStream_t st = net->ctx->stream;
  const int64_t N = net->N();
  const int nblk = 256;
  double* part = net->red;              // nb * nblk
  double* wbuf = net->red + (size_t)nb * nblk;  // nb
  hipLaunchKernelGGL(sumsq_kernel, dim3(nblk, nb), dim3(AB), 0, st, tgt, N, part);
  std::vector<double> hp((size_t)nb * nblk);
  CYX_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  std::vector<double> wb(nb);
  for (int b = 0; b < nb; ++b) {
    double s = 0;
    for (int k = 0; k < nblk; ++k) s += hp[(size_t)b * nblk + k];
    wb[b] = s > 0 ? 1.0 / (nb * s) : 0.0;
  }
  CYX_HIP(hipMemcpyAsync(wbuf, wb.data(), nb * sizeof(double), hipMemcpyHostToDevice, st));
  const float* sr = net->p + net->woff[NLAYER];
  float* ds = net->g + net->woff[NLAYER];
  hipLaunchKernelGGL(head_loss_kernel, dim3(nblk, nb), dim3(AB), 0, st, (const float*)A.f[Y], sr, tgt,
                     (const double*)wbuf, net->lv[0].nr, N, A.d[Y], ds, part);
  CYX_HIP(hipGetLastError());
  CYX_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  double l = 0;
  for (double v : hp) l += v;
  *loss = l;
  return CYX_OK;
}

int grads(cyx_unet_s* net, const float* x, const float* tgt, int nb, double* loss) {
  int rc;
  if ((rc = ensure_acts(net, nb)) != CYX_OK) return rc;
  hipStream_t st = net->ctx->stream;
  const Acts A = acts_of(net, nb);
  const int64_t N = net->N();
  CYX_HIP(hipMemcpyAsync(A.f[X], x, (size_t)nb * N * sizeof(float), hipMemcpyDeviceToDevice, st));
  if ((rc = forward(net, A, nb)) != CYX_OK) return rc;
  hipLaunchKernelGGL(zero_f_kernel, dim3(ceil_div(net->nparam, 256)), dim3(256), 0, st, net->g, net->nparam);
  if ((rc = head_loss(net, A, tgt, nb, loss)) != CYX_OK) return rc;
  return backward(net, A, nb);
}
}  // namespace

namespace cyx {
// u = U(D^{-1} r) for one FP64 residual r on the device (FP32 output); the
// kernels run on the network's context stream
int unet_precondition(cyx_unet_s* net, const double* coef, const double* r, float* u) {
  int rc;
  if ((rc = ensure_acts(net, 1)) != CYX_OK) return rc;
  hipStream_t st = net->ctx->stream;
  const Acts A = acts_of(net, 1);
  const int64_t N = net->N();
  hipLaunchKernelGGL(unet_input_kernel<double>, dim3(ceil_div(N, 256)), dim3(256), 0, st, r, coef, net->lv[0].nr, N,
                     1, (const float*)nullptr, A.f[X]);
  if ((rc = forward(net, A, 1)) != CYX_OK) return rc;
  hipLaunchKernelGGL(head_apply_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, (const float*)A.f[Y],
                     (const float*)(net->p + net->woff[NLAYER]), net->lv[0].nr, N, 1, u);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}
cyx_ctx_s* unet_context(const cyx_unet_s* net) { return net->ctx; }
bool unet_matches(const cyx_unet_s* net, const cyx_grid_desc& g) {
  return net->lv[0].nr == g.nr && net->lv[0].nt == g.nt && net->lv[0].nz == g.nz;
}
}  // namespace cyx

extern "C" {

int cyx_unet_create(cyx_ctx_t ctx, const cyx_grid_desc* grid, cyx_unet_t* out) {
  CYX_CHECK(ctx && grid && out, "null argument");
  CYX_CHECK(grid->nr % 4 == 0 && grid->nt % 4 == 0 && grid->nz % 4 == 0,
            "the U-Net needs nr, nt, nz divisible by 4 (got %d x %d x %d)", grid->nr, grid->nt, grid->nz);
  CYX_CHECK(grid->nr <= 512, "nr above 512 is not supported by the U-Net head");
  CYX_HIP(hipSetDevice(ctx->device));
  cyx_unet_s* n = new cyx_unet_s();
  n->ctx = ctx;
  n->lv[0] = Lvl{grid->nr, grid->nt, grid->nz};
  n->lv[1] = Lvl{grid->nr / 2, grid->nt / 2, grid->nz / 2};
  n->lv[2] = Lvl{grid->nr / 4, grid->nt / 4, grid->nz / 4};
  int64_t off = 0;
  for (int l = 0; l < NLAYER; ++l) {
    n->woff[l] = off;
    off += (int64_t)ntap(SPEC[l].mode) * SPEC[l].cin * SPEC[l].cout;
  }
  n->woff[NLAYER] = off;
  n->nparam = off + grid->nr;
  const size_t pb = (size_t)n->nparam * sizeof(float);
  bool ok = hipMalloc((void**)&n->p, pb) == hipSuccess && hipMalloc((void**)&n->g, pb) == hipSuccess &&
            hipMalloc((void**)&n->m1, pb) == hipSuccess && hipMalloc((void**)&n->m2, pb) == hipSuccess &&
            hipMalloc((void**)&n->wb, (size_t)off * 2) == hipSuccess &&
            hipMalloc((void**)&n->wt, (size_t)off * 2) == hipSuccess &&
            hipMalloc((void**)&n->red, (size_t)1 << 20) == hipSuccess;
  if (!ok) {
    (void)hipFree(n->p); (void)hipFree(n->g); (void)hipFree(n->m1); (void)hipFree(n->m2);
    (void)hipFree(n->wb); (void)hipFree(n->wt); (void)hipFree(n->red);
    delete n;
    set_error("U-Net allocation failed");
    return CYX_ERR_NOMEM;
  }
  CYX_HIP(hipMemset(n->p, 0, pb));
  CYX_HIP(hipMemset(n->m1, 0, pb));
  CYX_HIP(hipMemset(n->m2, 0, pb));
  *out = n;
  return CYX_OK;
}

int cyx_unet_destroy(cyx_unet_t n) {
  if (!n) return CYX_OK;
  (void)hipDeviceSynchronize();  // the context may already be gone
  (void)hipFree(n->p); (void)hipFree(n->g); (void)hipFree(n->m1); (void)hipFree(n->m2);
  (void)hipFree(n->wb); (void)hipFree(n->wt); (void)hipFree(n->red); (void)hipFree(n->act);
  delete n;
  return CYX_OK;
}

int64_t cyx_unet_num_params(cyx_unet_t n) { return n ? n->nparam : -1; }

int cyx_unet_set_params(cyx_unet_t n, const float* params, int reset_optimizer) {
  CYX_CHECK(n && params, "null argument");
  hipStream_t st = n->ctx->stream;
  CYX_HIP(hipMemcpyAsync(n->p, params, (size_t)n->nparam * sizeof(float), hipMemcpyHostToDevice, st));
  if (reset_optimizer) {
    CYX_HIP(hipMemsetAsync(n->m1, 0, (size_t)n->nparam * sizeof(float), st));
    CYX_HIP(hipMemsetAsync(n->m2, 0, (size_t)n->nparam * sizeof(float), st));
    n->adam_t = 0;
  }
  int rc = refresh_bf16(n);
  CYX_HIP(hipStreamSynchronize(st));
  return rc;
}

int cyx_unet_get_params(cyx_unet_t n, float* params) {
  CYX_CHECK(n && params, "null argument");
  CYX_HIP(hipMemcpyAsync(params, n->p, (size_t)n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
  CYX_HIP(hipStreamSynchronize(n->ctx->stream));
  return CYX_OK;
}

int cyx_unet_forward(cyx_unet_t n, const float* x, float* u, int nb) {
  CYX_CHECK(n && x && u && nb >= 1, "bad argument");
  int rc;
  if ((rc = ensure_acts(n, nb)) != CYX_OK) return rc;
  hipStream_t st = n->ctx->stream;
  const Acts A = acts_of(n, nb);
  const int64_t N = n->N();
  CYX_HIP(hipMemcpyAsync(A.f[X], x, (size_t)nb * N * sizeof(float), hipMemcpyDeviceToDevice, st));
  if ((rc = forward(n, A, nb)) != CYX_OK) return rc;
  hipLaunchKernelGGL(head_apply_kernel, dim3(ceil_div(N * nb, 256)), dim3(256), 0, st, (const float*)A.f[Y],
                     (const float*)(n->p + n->woff[NLAYER]), n->lv[0].nr, N, nb, u);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int cyx_unet_gradient(cyx_unet_t n, const float* x, const float* target, int nb, float* grad, double* loss) {
  CYX_CHECK(n && x && target && grad && loss && nb >= 1, "bad argument");
  int rc;
  if ((rc = grads(n, x, target, nb, loss)) != CYX_OK) return rc;
  CYX_HIP(hipMemcpyAsync(grad, n->g, (size_t)n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
  CYX_HIP(hipStreamSynchronize(n->ctx->stream));
  return CYX_OK;
}

int cyx_unet_train_step(cyx_unet_t n, const float* x, const float* target, int nb, float lr, double* loss) {
  CYX_CHECK(n && x && target && loss && nb >= 1 && lr > 0, "bad argument");
  int rc;
  if ((rc = grads(n, x, target, nb, loss)) != CYX_OK) return rc;
  n->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n->nparam, 256)), dim3(256), 0, n->ctx->stream, n->p,
                     (const float*)n->g, n->m1, n->m2, (int)n->nparam, lr, b1, b2, 1e-8f, c1, c2);
  CYX_HIP(hipGetLastError());
  return refresh_bf16(n);
}

int cyx_unet_input(cyx_unet_t n, const cyx_poisson_t ps, const double* r, const float* scale, int nb, float* x) {
  CYX_CHECK(n && ps && r && x && nb >= 1, "bad argument");
  const int64_t N = n->N();
  hipLaunchKernelGGL(unet_input_kernel<double>, dim3(ceil_div(N * nb, 256)), dim3(256), 0, n->ctx->stream, r,
                     (const double*)ps->coef, n->lv[0].nr, N, nb, scale, x);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

}  // extern "C"
