// Periodic-Pulay potential mixing on the device (src/mixing.m:20-116 with
// MixingPrecond = 0, PulayRestartFreq = 0) and the SCF error norm
// (src/scf.m:197-199).
//
// One mixing step, iteration `iter` (1-based):
//   f_k = g_k - x_k;  if iter > 1: X(:,i) = x_k - x_{k-1}, F(:,i) = f_k - f_{k-1}
//   Anderson (iter % p == 0 and iter > 1): Gamma = pinv(F'F) F'f_k,
//       x_w = x_k - X Gamma, f_w = f_k - F Gamma, x_{k+1} = x_w + beta f_w
//   else x_{k+1} = x_k + c f_k, c = beta when iter % p == 0 (iteration 1 with
//   p = 1), omega otherwise (src/mixing.m:99-103)
// Reductions are per-block partials plus a one-block finalisation in a
// fixed order (bitwise reproducible).
#include "cyx_internal.hpp"
#include "cyx_reduce.hpp"

namespace cyx {

struct MixState {
  double gamma[16];
};

// f = g - x; history columns; keep x, f as the previous iterate; Anderson dot partials
template <int MH>
__global__ void __launch_bounds__(AB) mix_hist_kernel(int64_t N, const double* __restrict__ g,
                                                      const double* __restrict__ x, double* __restrict__ f,
                                                      double* __restrict__ xkm1, double* __restrict__ fkm1,
                                                      double* __restrict__ X, double* __restrict__ F, int hcol,
                                                      int anderson, double* part) {
  __shared__ double red[AB];
  constexpr int ND = MH * (MH + 1) / 2 + MH;
  double acc[ND > 0 ? ND : 1];
#pragma unroll
  for (int d = 0; d < (ND > 0 ? ND : 1); ++d) acc[d] = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const double xk = x[n];
    const double fk = g[n] - xk;
    f[n] = fk;
    if (hcol >= 0) {
      X[(int64_t)hcol * N + n] = xk - xkm1[n];
      F[(int64_t)hcol * N + n] = fk - fkm1[n];
    }
    xkm1[n] = xk;
    fkm1[n] = fk;
    if constexpr (MH > 0) {
      if (anderson) {
        double fv[MH];
#pragma unroll
        for (int a = 0; a < MH; ++a) fv[a] = F[(int64_t)a * N + n];
        int d = 0;
#pragma unroll
        for (int a = 0; a < MH; ++a)
#pragma unroll
          for (int b = a; b < MH; ++b) acc[d++] += fv[a] * fv[b];
#pragma unroll
        for (int a = 0; a < MH; ++a) acc[d++] += fv[a] * fk;
      }
    }
  }
  if constexpr (MH > 0) {
    if (anderson) {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const double s = block_sum(acc[d], red);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * ND + d] = s;
      }
    }
  }
}

template <int MH>
__global__ void __launch_bounds__(AB) mix_solve_kernel(const double* part, int nblk, MixState* S) {
  constexpr int ND = MH * (MH + 1) / 2 + MH;
  __shared__ double tot[ND];
  reduce_partials(part, nblk, ND, tot);
  if (threadIdx.x == 0) {
    double A[MH * MH], b[MH], gm[MH];
    int d = 0;
#pragma unroll
    for (int a = 0; a < MH; ++a)
#pragma unroll
      for (int c = a; c < MH; ++c) {
        A[a * MH + c] = tot[d];
        A[c * MH + a] = tot[d];
        ++d;
      }
#pragma unroll
    for (int a = 0; a < MH; ++a) b[a] = tot[d++];
    pinv_solve<MH>(A, b, gm);
#pragma unroll
    for (int a = 0; a < MH; ++a) S->gamma[a] = gm[a];
  }
}

// x_out = x - X Gamma + c (f - F Gamma)   (MH = 0: x_out = x + c f)
template <int MH>
__global__ void __launch_bounds__(AB) mix_update_kernel(int64_t N, const double* __restrict__ x,
                                                        const double* __restrict__ f, const double* __restrict__ X,
                                                        const double* __restrict__ F, const MixState* S, double c,
                                                        double* __restrict__ xout) {
  double g[MH > 0 ? MH : 1];
#pragma unroll
  for (int k = 0; k < MH; ++k) g[k] = S->gamma[k];
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    double xw = x[n], fw = f[n];
#pragma unroll
    for (int k = 0; k < MH; ++k) {
      xw -= X[(int64_t)k * N + n] * g[k];
      fw -= F[(int64_t)k * N + n] * g[k];
    }
    xout[n] = xw + c * fw;
  }
}

// per-block partials of: Escc = sum (vout - vin) rho W (src/scf.m:162),
// ||vout - vin||^2 and ||vout||^2 (src/scf.m:184)
__global__ void __launch_bounds__(AB) scf_terms_kernel(int64_t N, const double* __restrict__ vout,
                                                       const double* __restrict__ vin,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ W, double* __restrict__ part) {
  __shared__ double red[AB];
  double e0 = 0, e1 = 0, e2 = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const double vo = vout[n];
    const double d = vo - vin[n];
    e0 += d * rho[n] * W[n];
    e1 += d * d;
    e2 += vo * vo;
  }
  e0 = block_sum(e0, red);
  e1 = block_sum(e1, red);
  e2 = block_sum(e2, red);
  if (threadIdx.x == 0) {
    part[3 * (int64_t)blockIdx.x + 0] = e0;
    part[3 * (int64_t)blockIdx.x + 1] = e1;
    part[3 * (int64_t)blockIdx.x + 2] = e2;
  }
}

}  // namespace cyx

struct cyx_mixer_s {
  cyx_ctx_s* ctx = nullptr;
  int64_t N = 0;
  int m = 7, p = 1;
  double beta = 0.3, omega = 0.3;
  double* buf = nullptr;  // X | F | xkm1 | fkm1 | f | partials | state
  int nblk = 0;
};

using namespace cyx;

template <int MH>
static int mix_step_impl(cyx_mixer_s* M, const double* g, const double* x, double* xout, int iter) {
  const int64_t N = M->N;
  hipStream_t st = M->ctx->stream;
  double* X = M->buf;
  double* F = X + (int64_t)MH * N;
  double* xkm1 = F + (int64_t)MH * N;
  double* fkm1 = xkm1 + N;
  double* f = fkm1 + N;
  double* part = f + N;
  MixState* S = (MixState*)(part + (int64_t)M->nblk * (MH * (MH + 1) / 2 + MH + 1));
  const int hcol = iter > 1 ? (iter - 2) % MH : -1;
  const bool anderson = (iter % M->p == 0) && iter > 1;
  const dim3 grid(M->nblk);
  if (anderson) {
    hipLaunchKernelGGL(mix_hist_kernel<MH>, grid, dim3(AB), 0, st, N, g, x, f, xkm1, fkm1, X, F, hcol, 1, part);
    hipLaunchKernelGGL(mix_solve_kernel<MH>, dim3(1), dim3(AB), 0, st, (const double*)part, M->nblk, S);
    hipLaunchKernelGGL(mix_update_kernel<MH>, grid, dim3(AB), 0, st, N, x, (const double*)f, (const double*)X,
                       (const double*)F, (const MixState*)S, M->beta, xout);
  } else {
    hipLaunchKernelGGL(mix_hist_kernel<0>, grid, dim3(AB), 0, st, N, g, x, f, xkm1, fkm1, X, F, hcol, 0, part);
    hipLaunchKernelGGL(mix_update_kernel<0>, grid, dim3(AB), 0, st, N, x, (const double*)f, (const double*)X,
                       (const double*)F, (const MixState*)S, (iter % M->p == 0) ? M->beta : M->omega, xout);
  }
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

extern "C" {

int cyx_mixer_create(cyx_ctx_t ctx, int64_t N, int history, double beta, double omega, int pulay_frequency,
                     const double* x0, cyx_mixer_t* out) {
  CYX_CHECK(ctx && out && x0 && N > 0, "null argument");
  CYX_CHECK(history >= 1 && history <= 8, "mixing history must be in [1, 8]");
  CYX_CHECK(pulay_frequency >= 1, "Pulay frequency must be >= 1");
  cyx_mixer_s* M = new cyx_mixer_s();
  M->ctx = ctx;
  M->N = N;
  M->m = history;
  M->p = pulay_frequency;
  M->beta = beta;
  M->omega = omega;
  M->nblk = (int)std::min<int64_t>(ceil_div(N, AB), 1024);
  const int nd = history * (history + 1) / 2 + history + 1;
  const size_t bytes = ((size_t)(2 * history + 3) * N + (size_t)M->nblk * nd) * sizeof(double) + sizeof(MixState) + 64;
  if (hipMalloc((void**)&M->buf, bytes) != hipSuccess) {
    delete M;
    set_error("mixer allocation of %zu bytes failed", bytes);
    return CYX_ERR_NOMEM;
  }
  // history zero (src/scf.m:42-44), x_{k-1} = x0, f_{k-1} = 0
  double* xkm1 = M->buf + (int64_t)2 * history * N;
  CYX_HIP(hipMemsetAsync(M->buf, 0, bytes, ctx->stream));
  CYX_HIP(hipMemcpyAsync(xkm1, x0, N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  *out = M;
  return CYX_OK;
}

int cyx_mixer_destroy(cyx_mixer_t M) {
  if (!M) return CYX_OK;
  // the context may already be gone: synchronise the device, not its stream
  (void)hipDeviceSynchronize();
  (void)hipFree(M->buf);
  delete M;
  return CYX_OK;
}

int cyx_mixer_step(cyx_mixer_t M, const double* g, const double* x, double* xout, int iter) {
  CYX_CHECK(M && g && x && xout && iter >= 1, "bad argument");
  switch (M->m) {
    case 1: return mix_step_impl<1>(M, g, x, xout, iter);
    case 2: return mix_step_impl<2>(M, g, x, xout, iter);
    case 3: return mix_step_impl<3>(M, g, x, xout, iter);
    case 4: return mix_step_impl<4>(M, g, x, xout, iter);
    case 5: return mix_step_impl<5>(M, g, x, xout, iter);
    case 6: return mix_step_impl<6>(M, g, x, xout, iter);
    case 7: return mix_step_impl<7>(M, g, x, xout, iter);
    default: return mix_step_impl<8>(M, g, x, xout, iter);
  }
}

int cyx_scf_terms(cyx_ham_t h, const double* vout, const double* vin, const double* rho, double out[3]) {
  CYX_CHECK(h && vout && vin && rho && out, "null argument");
  const int nblk = 512;
  int rc;
  if ((rc = ensure(h->red, 3 * (size_t)nblk * sizeof(double))) != CYX_OK) return rc;
  hipStream_t st = h->ctx->stream;
  hipLaunchKernelGGL(scf_terms_kernel, dim3(nblk), dim3(AB), 0, st, h->N(), vout, vin, rho, (const double*)h->W,
                     (double*)h->red.p);
  CYX_HIP(hipGetLastError());
  double hp[3 * 512];
  CYX_HIP(hipMemcpyAsync(hp, h->red.p, sizeof hp, hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  out[0] = out[1] = out[2] = 0;
  for (int k = 0; k < nblk; ++k)
    for (int j = 0; j < 3; ++j) out[j] += hp[3 * k + j];
  return CYX_OK;
}

}  // extern "C"
