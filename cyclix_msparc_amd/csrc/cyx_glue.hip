// SCF-step glue on the device, all FP64: electron density from the states
// (src/electronDensity.m:13-17), LDA (Perdew-Wang) exchange-correlation
// potential (src/exchangeCorrelationPotential.m:82-106), Veff = phi + Vxc
// (src/scf.m:39,155) and the grid sums of the total energy
// (src/evaluateTotalEnergy.m:27-52,72).
#include "cyx_internal.hpp"

namespace cyx {

// rho[n] += c * sum_s occ[s] |psi_s[n]|^2
template <class T>
__global__ void density_kernel(const T* __restrict__ psi, int64_t N, int64_t ld, int ns, const double* __restrict__ occ,
                               double c, double* __restrict__ rho) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double acc = 0;
  for (int s = 0; s < ns; ++s) {
    const T v = psi[(int64_t)s * ld + n];
    double a2;
    if constexpr (traits<T>::is_complex)
      a2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    else
      a2 = (double)v * (double)v;
    acc = __builtin_fma(occ[s], a2, acc);
  }
  rho[n] += c * acc;
}

struct PW {
  static constexpr double p = 1.0, A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  static constexpr double C2 = 0.73855876638202, C3 = 0.9847450218427;
};

__device__ inline double pw_rs(double rho) { return cbrt(0.75 / (M_PI * rho)); }

// Vxc(rho), src/exchangeCorrelationPotential.m:82-106
__global__ void lda_pw_kernel(int64_t N, const double* __restrict__ rho_in, const double* __restrict__ phi,
                              double* __restrict__ vxc, double* __restrict__ veff, double rhotol) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double rho = fmax(rho_in[n], rhotol);
  const double rs = pw_rs(rho);
  const double srs = sqrt(rs);
  const double G = 2 * PW::A * (PW::b1 * srs + PW::b2 * rs + PW::b3 * rs * srs + PW::b4 * rs * rs);
  const double dG = PW::A * (PW::b1 / srs + 2 * PW::b2 + 3 * PW::b3 * srs + 2 * (PW::p + 1) * PW::b4 * rs);
  const double lg = log(1 + 1 / G);
  const double ec = -2 * PW::A * (1 + PW::a1 * rs) * lg;
  const double vc = ec - (rs / 3) * (-2 * PW::A * PW::a1 * lg - (-2 * PW::A * (1 + PW::a1 * rs)) * dG / (G * G + G));
  const double v = -PW::C3 * cbrt(rho) + vc;
  vxc[n] = v;
  if (veff) veff[n] = phi[n] + v;
}

// per-block partial sums of: Exc (ec*rho*W - C2 rho^{4/3} W), Exc_dc (Vxc rho W), Eelec_dc (0.5 (b - rho) phi W)
__global__ void __launch_bounds__(256) energy_kernel(int64_t N, const double* __restrict__ rho_in,
                                                     const double* __restrict__ vxc, const double* __restrict__ phi,
                                                     const double* __restrict__ b, const double* __restrict__ W,
                                                     double rhotol, double* __restrict__ part) {
  __shared__ double red[3][256];
  double e0 = 0, e1 = 0, e2 = 0;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const double r_raw = rho_in[n];
    const double rho = fmax(r_raw, rhotol);
    const double rs = pw_rs(rho);
    const double srs = sqrt(rs);
    const double G = 2 * PW::A * (PW::b1 * srs + PW::b2 * rs + PW::b3 * rs * srs + PW::b4 * rs * rs);
    const double ec = -2 * PW::A * (1 + PW::a1 * rs) * log(1 + 1 / G);
    const double w = W[n];
    e0 += ec * rho * w - PW::C2 * rho * cbrt(rho) * w;
    e1 += vxc[n] * rho * w;
    e2 += 0.5 * (b[n] - r_raw) * phi[n] * w;
  }
  red[0][threadIdx.x] = e0;
  red[1][threadIdx.x] = e1;
  red[2][threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = red[k][0];
}

template <class R>
__global__ void cast_kernel(const double* __restrict__ src, R* __restrict__ dst, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) dst[n] = (R)src[n];
}

template <class T>
static int density_impl(cyx_ham_s* h, const void* psi, int ns, const double* occ_dev, double c, double* rho) {
  const int64_t N = h->N();
  hipLaunchKernelGGL(density_kernel<T>, dim3(ceil_div(N, 256)), dim3(256), 0, h->ctx->stream, (const T*)psi, N, N, ns,
                     occ_dev, c, rho);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int density_accumulate(cyx_ham_s* h, int dtype, const void* psi, int ns, const double* occ_dev, double c, double* rho) {
  switch (dtype) {
    case CYX_F64:
      return density_impl<double>(h, psi, ns, occ_dev, c, rho);
    case CYX_F32:
      return density_impl<float>(h, psi, ns, occ_dev, c, rho);
    case CYX_C128:
      return density_impl<cplx<double>>(h, psi, ns, occ_dev, c, rho);
    case CYX_C64:
      return density_impl<cplx<float>>(h, psi, ns, occ_dev, c, rho);
  }
  return CYX_ERR_ARG;
}

int lda_pw(hipStream_t st, int64_t N, const double* rho, const double* phi, double* vxc, double* veff, double rhotol) {
  hipLaunchKernelGGL(lda_pw_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, rho, phi, vxc, veff, rhotol);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int energy_terms(hipStream_t st, int64_t N, const double* rho, const double* vxc, const double* phi, const double* b,
                 const double* W, double rhotol, double* part_dev, int nblk, double out[3]) {
  hipLaunchKernelGGL(energy_kernel, dim3(nblk), dim3(256), 0, st, N, rho, vxc, phi, b, W, rhotol, part_dev);
  CYX_HIP(hipGetLastError());
  std::vector<double> hp(3 * (size_t)nblk);
  CYX_HIP(hipMemcpyAsync(hp.data(), part_dev, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  out[0] = out[1] = out[2] = 0;
  for (int k = 0; k < nblk; ++k)
    for (int j = 0; j < 3; ++j) out[j] += hp[3 * k + j];
  return CYX_OK;
}

int veff_from_device(cyx_ham_s* h, const double* veff) {
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  CYX_HIP(hipMemcpyAsync(h->c64.veff, veff, N * sizeof(double), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(cast_kernel<float>, dim3(ceil_div(N, 256)), dim3(256), 0, st, veff, h->c32.veff, N);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

}  // namespace cyx
