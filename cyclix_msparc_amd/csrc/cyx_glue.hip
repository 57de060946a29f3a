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

// drho[n] += c * sum_s (dg[s] |psi_s[n]|^2 + 2 occ[s] Re(conj(psi_s[n]) dpsi_s[n]))
// (src/phonon.m:1236-1238, q = 0); coef = {occ[0..ns), dg[0..ns)}
template <class T>
__global__ void density_response_kernel(const T* __restrict__ psi, const T* __restrict__ dpsi, int64_t N, int64_t ld,
                                        int ns, const double* __restrict__ coef, double c,
                                        double* __restrict__ drho) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double acc = 0;
  for (int s = 0; s < ns; ++s) {
    const T v = psi[(int64_t)s * ld + n];
    const T d = dpsi[(int64_t)s * ld + n];
    double a2, re;
    if constexpr (traits<T>::is_complex) {
      a2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
      re = (double)v.x * (double)d.x + (double)v.y * (double)d.y;
    } else {
      a2 = (double)v * (double)v;
      re = (double)v * (double)d;
    }
    acc = __builtin_fma(coef[ns + s], a2, acc);
    acc = __builtin_fma(2.0 * coef[s], re, acc);
  }
  drho[n] += c * acc;
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

// dVxc/drho of the potential above (the DFPT kernel S.Dvxc_drho,
// src/phonon.m:215,303): PW92 differentiated in closed form through
// rs = (3/(4 pi rho))^(1/3): dvc/drho = (2 ec'/3 - rs ec''/3)(-rs/(3 rho)),
// dvx/drho = -C3 rho^(-2/3)/3; 0 where rho is clamped at rhotol.  Outputs
// (each optional): fxc; dveff = dphi + fxc drho (src/phonon.m:303-306).
__global__ void lda_pw_fxc_kernel(int64_t N, const double* __restrict__ rho_in, const double* __restrict__ drho,
                                  const double* __restrict__ dphi, double* __restrict__ fxc,
                                  double* __restrict__ dveff, double rhotol) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double r0 = rho_in[n];
  double f = 0;
  if (r0 >= rhotol) {
    const double rs = pw_rs(r0);
    const double srs = sqrt(rs);
    const double G = 2 * PW::A * (PW::b1 * srs + PW::b2 * rs + PW::b3 * rs * srs + PW::b4 * rs * rs);
    const double G1 = PW::A * (PW::b1 / srs + 2 * PW::b2 + 3 * PW::b3 * srs + 2 * (PW::p + 1) * PW::b4 * rs);
    const double G2 = PW::A * (-0.5 * PW::b1 / (rs * srs) + 1.5 * PW::b3 / srs + 2 * PW::p * (PW::p + 1) * PW::b4);
    const double D = G * G + G;
    const double lg = log(1 + 1 / G);
    const double ec1 = -2 * PW::A * PW::a1 * lg + 2 * PW::A * (1 + PW::a1 * rs) * G1 / D;
    const double ec2 = 4 * PW::A * PW::a1 * G1 / D + 2 * PW::A * (1 + PW::a1 * rs) * (G2 * D - G1 * G1 * (2 * G + 1)) / (D * D);
    const double cr = cbrt(r0);
    f = (2 * ec1 / 3 - rs * ec2 / 3) * (-rs / (3 * r0)) - PW::C3 / (3 * cr * cr);
  }
  if (fxc) fxc[n] = f;
  if (dveff) dveff[n] = dphi[n] + f * drho[n];
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

template <class T>
static int density_response_impl(cyx_ham_s* h, const void* psi, const void* dpsi, int ns, const double* coef_dev,
                                 double c, double* drho) {
  const int64_t N = h->N();
  hipLaunchKernelGGL(density_response_kernel<T>, dim3(ceil_div(N, 256)), dim3(256), 0, h->ctx->stream, (const T*)psi,
                     (const T*)dpsi, N, N, ns, coef_dev, c, drho);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int density_response(cyx_ham_s* h, int dtype, const void* psi, const void* dpsi, int ns, const double* coef_dev,
                     double c, double* drho) {
  switch (dtype) {
    case CYX_F64:
      return density_response_impl<double>(h, psi, dpsi, ns, coef_dev, c, drho);
    case CYX_F32:
      return density_response_impl<float>(h, psi, dpsi, ns, coef_dev, c, drho);
    case CYX_C128:
      return density_response_impl<cplx<double>>(h, psi, dpsi, ns, coef_dev, c, drho);
    case CYX_C64:
      return density_response_impl<cplx<float>>(h, psi, dpsi, ns, coef_dev, c, drho);
  }
  return CYX_ERR_ARG;
}

int lda_pw_fxc(hipStream_t st, int64_t N, const double* rho, const double* drho, const double* dphi, double* fxc,
               double* dveff, double rhotol) {
  hipLaunchKernelGGL(lda_pw_fxc_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, rho, drho, dphi, fxc, dveff,
                     rhotol);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
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
