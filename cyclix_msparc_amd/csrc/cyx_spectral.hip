// Exact inverse of S.Lap_std (k = 0) on a Cyclix grid.
//
// The reference preconditions its Poisson AAR (src/poissonSolve.m:36) and
// its Sternheimer BiCGSTAB (src/phonon.m:1519) with ILU factors of Lap_std
// (src/initialization.m:752, droptol 1e-5), which sequential triangular
// solves apply.  Here the same operator is inverted exactly and in parallel:
// Lap_std = (DL11 + DG1) (x) I + diag(R2inv) (x) DL22 + DL33 [+ helical:
// DL22 + DG2 (x) DG3] (src/initialization.m:734-741), whose theta and z
// factors (src/blochLaplacian_1d.m at k = 0) are circulant.  A 2-D Fourier
// transform over (theta, z) therefore turns it into one real nr x nr band
// matrix per mode (m, n):
//
//   A_mn = (DL11 + DG1) + lam22(m) diag(R2inv) + lam33(n) I
//          [+ (lam22(m) + lamG2(m) lamG3(n)) I],
//
// lam the circulant symbols: lam22(m) = (w2_0 + 2 sum_p w2_p cos(2 pi p m/nt))
// tau^2/dt^2 (tau^2 only for helical cells, where R2inv carries 1/tau^2),
// lam33(n) the same in z over dz^2, and lamG2 lamG3 = 8 tau S_t(m) S_z(n) /
// (dt dz) with S(x) = sum_p w1_p sin(p x).  The band matrices are factored
// once (LU without pivoting: Lap_std is negative definite and self-adjoint
// in the metric r dr dt dz, so each A_mn is similar to a symmetric negative
// definite matrix and every pivot is negative).  A solve is a rocFFT forward
// transform into mode-major order, the band forward/back substitutions (one
// thread per mode, reading the factors mode-fastest), and the inverse
// transform.
#include <rocfft/rocfft.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "cyx_internal.hpp"

namespace cyx {

constexpr int BW = FDN;            // band half-width of A_mn
constexpr int NBD = 2 * BW + 1;    // stored diagonals per row: L (BW), U (1 + BW)

struct Spectral {
  int nr = 0, nt = 0, nz = 0, nmode = 0;
  int64_t N = 0;
  double* lu = nullptr;       // [nr][NBD][nmode]: l_{i,i-BW+d} (d < BW), 1/u_ii (d = BW), u_{i,i+d-BW}
  cplx<double>* buf = nullptr;  // two complex work vectors per column
  int buf_cols = 0;
  rocfft_plan fwd = nullptr, inv = nullptr;  // one column each
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  size_t work_bytes = 0;
};

static std::once_flag g_rocfft_once;
static rocfft_status g_rocfft_setup = rocfft_status_success;

#define CYX_FFT(call)                                                       \
  do {                                                                      \
    const rocfft_status s_ = (call);                                        \
    if (s_ != rocfft_status_success) {                                      \
      ::cyx::set_error("%s:%d %s: rocFFT status %d", __FILE__, __LINE__, #call, (int)s_); \
      return CYX_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

// Band LU of every mode on the host; returns false if a pivot is not
// negative (the operator would not be the negative definite Lap_std).
static bool factor_modes(const cyx_grid_desc& g, std::vector<double>& lu) {
  double w2[FDN + 1], w1[FDN + 1];
  host_fd_weights(w2, w1);
  const int nr = g.nr, nt = g.nt, nz = g.nz, nmode = nt * nz;
  const bool hel = g.cell_typ == 4 || g.cell_typ == 5;
  const double tau = hel ? g.twist : 0.0;
  const double pi2 = 2.0 * M_PI;
  std::vector<double> l22(nt), l33(nz), st(nt), sz(nz);
  for (int m = 0; m < nt; ++m) {
    double c = w2[0], s = 0.0;
    for (int p = 1; p <= FDN; ++p) {
      c += 2.0 * w2[p] * std::cos(pi2 * p * m / nt);
      s += w1[p] * std::sin(pi2 * p * m / nt);
    }
    l22[m] = c / (g.dt * g.dt);  // theta second-derivative symbol (without tau^2 or 1/r^2)
    st[m] = s;
  }
  for (int n = 0; n < nz; ++n) {
    double c = w2[0], s = 0.0;
    for (int p = 1; p <= FDN; ++p) {
      c += 2.0 * w2[p] * std::cos(pi2 * p * n / nz);
      s += w1[p] * std::sin(pi2 * p * n / nz);
    }
    l33[n] = c / (g.dz * g.dz);
    sz[n] = s;
  }
  lu.assign((size_t)nr * NBD * nmode, 0.0);
  std::vector<double> a((size_t)nr * NBD);  // row i: entries at columns i-BW .. i+BW
  for (int n = 0; n < nz; ++n) {
    for (int m = 0; m < nt; ++m) {
      const int mode = m + nt * n;
      const double shift = l33[n] + (hel ? tau * tau * l22[m] + 8.0 * tau * st[m] * sz[n] / (g.dt * g.dz) : 0.0);
      for (int i = 0; i < nr; ++i) {
        const double r = g.xin + i * g.dr;
        double* ai = &a[(size_t)i * NBD];
        for (int d = 0; d < NBD; ++d) ai[d] = 0.0;
        for (int p = 1; p <= FDN; ++p) {
          if (i + p < nr) ai[BW + p] = w2[p] / (g.dr * g.dr) + w1[p] / (r * g.dr);
          if (i - p >= 0) ai[BW - p] = w2[p] / (g.dr * g.dr) - w1[p] / (r * g.dr);
        }
        ai[BW] = w2[0] / (g.dr * g.dr) + l22[m] / (r * r) + shift;
      }
      // banded Gaussian elimination without pivoting: row i's entry at
      // column i-BW+d is a[i][d]
      for (int k = 0; k < nr; ++k) {
        const double piv = a[(size_t)k * NBD + BW];
        if (!(piv < 0.0)) return false;
        for (int i = k + 1; i <= std::min(nr - 1, k + BW); ++i) {
          double* ai = &a[(size_t)i * NBD];
          const int dk = k - i + BW;  // column k in row i
          const double l = ai[dk] / piv;
          ai[dk] = l;
          for (int j = k + 1; j <= std::min(nr - 1, k + BW); ++j) ai[j - i + BW] -= l * a[(size_t)k * NBD + (j - k + BW)];
        }
      }
      for (int i = 0; i < nr; ++i)
        for (int d = 0; d < NBD; ++d) {
          double v = a[(size_t)i * NBD + d];
          if (d == BW) v = 1.0 / v;
          lu[((size_t)i * NBD + d) * nmode + mode] = v;
        }
    }
  }
  return true;
}

// rocFFT plans over one column of length N: the forward transform reads the
// grid order [k][j][i] (j stride nr, k stride nr nt, the nr radii batched at
// distance 1) and writes the mode-major order [i][k][j] (mode m + nt n
// contiguous per radius); the inverse transform goes back
static int make_plans(Spectral* S) {
  const size_t nr = S->nr, nt = S->nt, nz = S->nz;
  const size_t len[2] = {nt, nz};  // fastest first
  const size_t grid_str[2] = {nr, nr * nt};
  const size_t mode_str[2] = {1, nt};
  for (int dir = 0; dir < 2; ++dir) {
    rocfft_plan_description d = nullptr;
    CYX_FFT(rocfft_plan_description_create(&d));
    const bool f = dir == 0;
    const rocfft_status s = rocfft_plan_description_set_data_layout(
        d, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr, nullptr, 2,
        f ? grid_str : mode_str, f ? 1 : nt * nz, 2, f ? mode_str : grid_str, f ? nt * nz : 1);
    if (s != rocfft_status_success) {
      (void)rocfft_plan_description_destroy(d);
      set_error("rocFFT data layout: status %d", (int)s);
      return CYX_ERR_HIP;
    }
    rocfft_plan* pl = f ? &S->fwd : &S->inv;
    const rocfft_status s2 =
        rocfft_plan_create(pl, rocfft_placement_notinplace,
                           f ? rocfft_transform_type_complex_forward : rocfft_transform_type_complex_inverse,
                           rocfft_precision_double, 2, len, nr, d);
    (void)rocfft_plan_description_destroy(d);
    if (s2 != rocfft_status_success) {
      set_error("rocFFT plan (%zu x %zu, batch %zu): status %d", nt, nz, nr, (int)s2);
      return CYX_ERR_HIP;
    }
    size_t wb = 0;
    CYX_FFT(rocfft_plan_get_work_buffer_size(*pl, &wb));
    if (wb > S->work_bytes) {
      if (S->work) (void)hipFree(S->work);
      S->work = nullptr;
      S->work_bytes = 0;
      CYX_HIP(hipMalloc(&S->work, wb));
      S->work_bytes = wb;
    }
  }
  return CYX_OK;
}

int spectral_create(const cyx_grid_desc& g, Spectral** out) {
  *out = nullptr;
  std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
  CYX_CHECK(g_rocfft_setup == rocfft_status_success, "rocfft_setup failed (%d)", (int)g_rocfft_setup);
  std::vector<double> lu;
  CYX_CHECK(factor_modes(g, lu), "Lap_std is not negative definite on this grid (a band pivot is >= 0)");
  Spectral* S = new Spectral();
  S->nr = g.nr;
  S->nt = g.nt;
  S->nz = g.nz;
  S->nmode = g.nt * g.nz;
  S->N = (int64_t)g.nr * g.nt * g.nz;
  int rc = CYX_OK;
  if (hipMalloc((void**)&S->lu, lu.size() * sizeof(double)) != hipSuccess ||
      hipMemcpy(S->lu, lu.data(), lu.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("spectral factors: device allocation of %zu bytes failed", lu.size() * sizeof(double));
    rc = CYX_ERR_NOMEM;
  }
  if (rc == CYX_OK && rocfft_execution_info_create(&S->info) != rocfft_status_success) {
    set_error("rocfft_execution_info_create failed");
    rc = CYX_ERR_HIP;
  }
  if (rc == CYX_OK) rc = make_plans(S);
  if (rc != CYX_OK) {
    spectral_destroy(S);
    return rc;
  }
  *out = S;
  return CYX_OK;
}

void spectral_destroy(Spectral* S) {
  if (!S) return;
  if (S->fwd) (void)rocfft_plan_destroy(S->fwd);
  if (S->inv) (void)rocfft_plan_destroy(S->inv);
  if (S->info) (void)rocfft_execution_info_destroy(S->info);
  (void)hipFree(S->work);
  (void)hipFree(S->lu);
  (void)hipFree(S->buf);
  delete S;
}


// forward and back substitution of every mode, in place on the mode-major
// spectrum Y ([i][mode], ncol columns of nr x nmode), scaled by `scale`.
// One thread per mode; the rows go in chunks of BSU whose factor entries and
// right-hand sides are all loaded before the chunk's recurrence, so a thread
// waits on memory once per chunk instead of once per row.
constexpr int BSU = 8;
constexpr int BSB = 64;  // threads per block: more blocks for few modes
__global__ void __launch_bounds__(BSB) band_solve_kernel(const double* __restrict__ lu, cplx<double>* __restrict__ Y,
                                                         int nr, int nmode, double scale) {
  const int mode = blockIdx.x * BSB + threadIdx.x;
  if (mode >= nmode) return;
  cplx<double>* y = Y + (int64_t)blockIdx.y * nr * nmode + mode;
  const int64_t rs = nmode;                  // stride between radii
  const int64_t ds = nmode;                  // stride between stored diagonals
  const int64_t ls = (int64_t)NBD * nmode;   // stride between rows of the factors
  const double* lm = lu + mode;
  // forward: y_i -= sum_{d=1..BW} l_{i,i-d} y_{i-d}; w[BW - d] holds y_{i-d}
  cplx<double> w[BW];
#pragma unroll
  for (int d = 0; d < BW; ++d) w[d] = cplx<double>{0.0, 0.0};
  for (int i0 = 0; i0 < nr; i0 += BSU) {
    double c[BSU][BW];
    cplx<double> b[BSU];
#pragma unroll
    for (int u = 0; u < BSU; ++u) {
      const int i = i0 + u;
      b[u] = i < nr ? y[i * rs] : cplx<double>{0.0, 0.0};
#pragma unroll
      for (int d = 1; d <= BW; ++d) c[u][d - 1] = i < nr ? lm[i * ls + (BW - d) * ds] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < BSU; ++u) {
      cplx<double> v = b[u];
#pragma unroll
      for (int d = 1; d <= BW; ++d) {
        v.x -= c[u][d - 1] * w[BW - d].x;
        v.y -= c[u][d - 1] * w[BW - d].y;
      }
#pragma unroll
      for (int d = 0; d < BW - 1; ++d) w[d] = w[d + 1];
      w[BW - 1] = v;
      if (i0 + u < nr) y[(i0 + u) * rs] = v;
    }
  }
  // back: x_i = (y_i - sum_{d=1..BW} u_{i,i+d} x_{i+d}) / u_ii; w[d - 1] holds x_{i+d}
#pragma unroll
  for (int d = 0; d < BW; ++d) w[d] = cplx<double>{0.0, 0.0};
  for (int i0 = nr - 1; i0 >= 0; i0 -= BSU) {
    double c[BSU][BW], inv[BSU];
    cplx<double> b[BSU];
#pragma unroll
    for (int u = 0; u < BSU; ++u) {
      const int i = i0 - u;
      b[u] = i >= 0 ? y[i * rs] : cplx<double>{0.0, 0.0};
      inv[u] = i >= 0 ? lm[i * ls + BW * ds] : 0.0;
#pragma unroll
      for (int d = 1; d <= BW; ++d) c[u][d - 1] = i >= 0 ? lm[i * ls + (BW + d) * ds] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < BSU; ++u) {
      cplx<double> v = b[u];
#pragma unroll
      for (int d = 1; d <= BW; ++d) {
        v.x -= c[u][d - 1] * w[d - 1].x;
        v.y -= c[u][d - 1] * w[d - 1].y;
      }
      v.x *= inv[u];
      v.y *= inv[u];
#pragma unroll
      for (int d = BW - 1; d > 0; --d) w[d] = w[d - 1];
      w[0] = v;
      if (i0 - u >= 0) y[(i0 - u) * rs] = cplx<double>{v.x * scale, v.y * scale};
    }
  }
}

// real r (ncol columns, ld) -> complex work column
__global__ void widen_kernel(const double* __restrict__ r, int64_t ld, cplx<double>* __restrict__ z, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n < N) z[(int64_t)blockIdx.y * N + n] = cplx<double>{r[(int64_t)blockIdx.y * ld + n], 0.0};
}
__global__ void real_part_kernel(const cplx<double>* __restrict__ z, double* __restrict__ x, int64_t ld, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n < N) x[(int64_t)blockIdx.y * ld + n] = z[(int64_t)blockIdx.y * N + n].x;
}

static int ensure_buf(Spectral* S, int ncol) {
  if (S->buf_cols >= ncol) return CYX_OK;
  (void)hipFree(S->buf);
  S->buf = nullptr;
  S->buf_cols = 0;
  if (hipMalloc((void**)&S->buf, 2 * (size_t)ncol * S->N * sizeof(cplx<double>)) != hipSuccess) {
    set_error("spectral work buffer of %d columns: allocation failed", ncol);
    return CYX_ERR_NOMEM;
  }
  S->buf_cols = ncol;
  return CYX_OK;
}

// Z = Lap_std^-1 R for complex columns with leading dimension N; the input
// columns are overwritten (rocFFT may use its input as scratch)
static int solve_inplace_complex(Spectral* S, cplx<double>* R, cplx<double>* mid, int ncol, hipStream_t st) {
  CYX_FFT(rocfft_execution_info_set_stream(S->info, st));
  if (S->work_bytes > 0) CYX_FFT(rocfft_execution_info_set_work_buffer(S->info, S->work, S->work_bytes));
  const double scale = 1.0 / ((double)S->nt * S->nz);
  for (int c = 0; c < ncol; ++c) {
    void* in[1] = {R + (int64_t)c * S->N};
    void* out[1] = {mid + (int64_t)c * S->N};
    CYX_FFT(rocfft_execute(S->fwd, in, out, S->info));
  }
  hipLaunchKernelGGL(band_solve_kernel, dim3(ceil_div(S->nmode, BSB), ncol), dim3(BSB), 0, st, (const double*)S->lu,
                     mid, S->nr, S->nmode, scale);
  CYX_HIP(hipGetLastError());
  for (int c = 0; c < ncol; ++c) {
    void* in[1] = {mid + (int64_t)c * S->N};
    void* out[1] = {R + (int64_t)c * S->N};
    CYX_FFT(rocfft_execute(S->inv, in, out, S->info));
  }
  return CYX_OK;
}

int spectral_solve_real(Spectral* S, const double* r, double* z, int ncol, int64_t ld, hipStream_t st) {
  int rc;
  if ((rc = ensure_buf(S, ncol)) != CYX_OK) return rc;
  cplx<double>* a = S->buf;
  cplx<double>* b = S->buf + (int64_t)ncol * S->N;
  const dim3 grid(ceil_div(S->N, 256), ncol);
  hipLaunchKernelGGL(widen_kernel, grid, dim3(256), 0, st, r, ld, a, S->N);
  if ((rc = solve_inplace_complex(S, a, b, ncol, st)) != CYX_OK) return rc;
  hipLaunchKernelGGL(real_part_kernel, grid, dim3(256), 0, st, (const cplx<double>*)a, z, ld, S->N);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int spectral_solve_complex(Spectral* S, const cplx<double>* r, cplx<double>* z, int ncol, int64_t ld, hipStream_t st) {
  int rc;
  if ((rc = ensure_buf(S, ncol)) != CYX_OK) return rc;
  cplx<double>* a = S->buf;
  cplx<double>* b = S->buf + (int64_t)ncol * S->N;
  CYX_HIP(hipMemcpy2DAsync(a, S->N * sizeof(cplx<double>), r, ld * sizeof(cplx<double>), S->N * sizeof(cplx<double>),
                           ncol, hipMemcpyDeviceToDevice, st));
  if ((rc = solve_inplace_complex(S, a, b, ncol, st)) != CYX_OK) return rc;
  CYX_HIP(hipMemcpy2DAsync(z, ld * sizeof(cplx<double>), a, S->N * sizeof(cplx<double>), S->N * sizeof(cplx<double>),
                           ncol, hipMemcpyDeviceToDevice, st));
  return CYX_OK;
}

}  // namespace cyx
