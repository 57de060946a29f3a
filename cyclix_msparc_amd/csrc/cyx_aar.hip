// Jacobi-preconditioned Alternating Anderson-Richardson Poisson solve
// (src/aar.m:33-72, src/poissonSolve.m:17-51; Jacobi preconditioner of
// extras/AAR_old.m:45-46,73-75).  Operator: S.Lap_std at k = 0
// (src/initialization.m:734-741), applied matrix-free.
//
// The iteration runs without host round trips: every kernel of iteration
// `count` first reads a device-side `done` flag and returns at once when an
// earlier iteration already met the tolerance.  Reductions use per-block
// partials and a last-block-done finalisation, summed in a fixed order, so a
// run is bitwise reproducible.  The host enqueues iterations in chunks and
// polls the flag between chunks.
#include "cyx_internal.hpp"

namespace cyx {

struct AarState {
  int done;       // 1 once relres <= tol
  int iters;      // count at exit (as src/aar.m reports count-1 after the loop)
  int final_buf;  // which x buffer holds the returned iterate
  int counter;    // last-block-done ticket
  double relres;  // relres of the returned iterate
  double nb_inv;  // 1 / ||f||
  double gamma[16];
};

constexpr int AB = 256;
constexpr int MAXHIST = 15;

struct PoissonOp {
  int nr, nt, nz;
  const double* coef;  // [nr][RCOEF]: c+[6], c-[6], c_theta, diag
  double wt2[FDN + 1];  // w2_p (theta second derivative, times c_theta per radius)
  double wz2[FDN + 1];  // w2_q / dz^2
  double cx[FDN + 1][FDN + 1];  // (-2 tau w1_p / dt) (w1_q / dz)
  int helical;
};

__device__ inline double lap_at(const PoissonOp& P, const double* __restrict__ x, int i, int j, int k) {
  const int nrt = P.nr * P.nt;
  const double* cf = P.coef + i * RCOEF;
  const int64_t n = (int64_t)k * nrt + j * P.nr + i;
  double acc = cf[2 * FDN + 1] * x[n];
#pragma unroll
  for (int p = 1; p <= FDN; ++p) {
    if (i + p < P.nr) acc = __builtin_fma(cf[p - 1], x[n + p], acc);
    if (i - p >= 0) acc = __builtin_fma(cf[FDN + p - 1], x[n - p], acc);
  }
  double st = 0.0;
#pragma unroll
  for (int p = 1; p <= FDN; ++p) {
    int jp = j + p;
    if (jp >= P.nt) jp -= P.nt;
    int jm = j - p;
    if (jm < 0) jm += P.nt;
    st = __builtin_fma(P.wt2[p], x[(int64_t)k * nrt + jp * P.nr + i] + x[(int64_t)k * nrt + jm * P.nr + i], st);
  }
  acc = __builtin_fma(cf[2 * FDN], st, acc);
#pragma unroll
  for (int q = 1; q <= FDN; ++q) {
    int kp = k + q;
    if (kp >= P.nz) kp -= P.nz;
    int km = k - q;
    if (km < 0) km += P.nz;
    acc = __builtin_fma(P.wz2[q], x[(int64_t)kp * nrt + j * P.nr + i] + x[(int64_t)km * nrt + j * P.nr + i], acc);
  }
  if (P.helical) {
#pragma unroll
    for (int q = 1; q <= FDN; ++q) {
      int kp = k + q;
      if (kp >= P.nz) kp -= P.nz;
      int km = k - q;
      if (km < 0) km += P.nz;
#pragma unroll
      for (int p = 1; p <= FDN; ++p) {
        int jp = j + p;
        if (jp >= P.nt) jp -= P.nt;
        int jm = j - p;
        if (jm < 0) jm += P.nt;
        const double d = (x[(int64_t)kp * nrt + jp * P.nr + i] - x[(int64_t)kp * nrt + jm * P.nr + i]) -
                         (x[(int64_t)km * nrt + jp * P.nr + i] - x[(int64_t)km * nrt + jm * P.nr + i]);
        acc = __builtin_fma(P.cx[p][q], d, acc);
      }
    }
  }
  return acc;
}

__device__ inline double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = AB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

// last-block-done: returns true in exactly one block (the last to arrive)
__device__ inline bool last_block(AarState* S) {
  __shared__ int ticket;
  __threadfence();
  if (threadIdx.x == 0) ticket = atomicAdd(&S->counter, 1);
  __syncthreads();
  const bool last = ticket == (int)gridDim.x - 1;
  if (last) __threadfence();
  return last;
}

// ||f|| -> nb_inv
__global__ void __launch_bounds__(AB) norm_kernel(const double* __restrict__ f, int64_t N, double* part, AarState* S) {
  __shared__ double red[AB];
  double a = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) a += f[n] * f[n];
  const double s = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (last_block(S)) {
    if (threadIdx.x == 0) {
      double t = 0;
      for (int b = 0; b < (int)gridDim.x; ++b) t += part[b];
      S->nb_inv = 1.0 / sqrt(t);
      S->counter = 0;
    }
  }
}

// res1 = f - A x ; relres ; res = res1 ./ diag
__global__ void __launch_bounds__(AB) residual_kernel(PoissonOp P, const double* __restrict__ f,
                                                      const double* __restrict__ x, double* __restrict__ res,
                                                      double* part, AarState* S, int count, int max_iter, double tol,
                                                      int xbuf) {
  if (S->done) return;
  __shared__ double red[AB];
  const int64_t N = (int64_t)P.nr * P.nt * P.nz;
  double a = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const int i = (int)(n % P.nr);
    const int j = (int)((n / P.nr) % P.nt);
    const int k = (int)(n / ((int64_t)P.nr * P.nt));
    const double r1 = f[n] - lap_at(P, x, i, j, k);
    a += r1 * r1;
    res[n] = r1 / P.coef[i * RCOEF + 2 * FDN + 1];
  }
  const double s = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (last_block(S)) {
    if (threadIdx.x == 0) {
      double t = 0;
      for (int b = 0; b < (int)gridDim.x; ++b) t += part[b];
      const double relres = sqrt(t) * S->nb_inv;
      S->relres = relres;
      S->iters = count;
      S->final_buf = xbuf;
      if (relres <= tol || count >= max_iter) S->done = (relres <= tol) ? 1 : 2;
      S->counter = 0;
    }
  }
}

// history update + Richardson step (count % p != 0)
__global__ void __launch_bounds__(AB) richardson_kernel(int64_t N, const double* __restrict__ xp,
                                                        const double* __restrict__ res, const double* __restrict__ fold,
                                                        double* xnew_xold, double* DX, double* DF, int hcol,
                                                        double omega, const AarState* S) {
  if (S->done) return;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const double x = xp[n];
    const double r = res[n];
    if (hcol >= 0) {
      DX[(int64_t)hcol * N + n] = x - xnew_xold[n];
      DF[(int64_t)hcol * N + n] = r - fold[n];
    }
    xnew_xold[n] = x + omega * r;
  }
}

// symmetric m x m pseudo-inverse times vector, MATLAB pinv tolerance
// max(size) * norm * eps, by cyclic Jacobi eigen-decomposition (one thread)
__device__ void pinv_solve(int m, double* A /*m*m*/, const double* b, double* out) {
  double V[MAXHIST * MAXHIST];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) V[i * m + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        tot += A[i * m + j] * A[i * m + j];
        if (i != j) off += A[i * m + j] * A[i * m + j];
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[p * m + q];
        if (apq == 0.0) continue;
        const double app = A[p * m + p], aqq = A[q * m + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = A[k * m + p], akq = A[k * m + q];
          A[k * m + p] = c * akp - s * akq;
          A[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = A[p * m + k], aqk = A[q * m + k];
          A[p * m + k] = c * apk - s * aqk;
          A[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = V[k * m + p], vkq = V[k * m + q];
          V[k * m + p] = c * vkp - s * vkq;
          V[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  double smax = 0;
  for (int i = 0; i < m; ++i) smax = fmax(smax, fabs(A[i * m + i]));
  const double tol = m * smax * 2.220446049250313e-16;
  for (int i = 0; i < m; ++i) out[i] = 0;
  for (int e = 0; e < m; ++e) {
    const double lam = A[e * m + e];
    if (fabs(lam) <= tol) continue;
    double vb = 0;
    for (int k = 0; k < m; ++k) vb += V[k * m + e] * b[k];
    vb /= lam;
    for (int k = 0; k < m; ++k) out[k] += V[k * m + e] * vb;
  }
}

// history update + DF^T DF, DF^T res partials; last block solves gamma
__global__ void __launch_bounds__(AB) anderson_dots_kernel(int64_t N, const double* __restrict__ xp,
                                                           const double* __restrict__ res,
                                                           const double* __restrict__ fold,
                                                           const double* __restrict__ xold, double* DX, double* DF,
                                                           int hcol, int m, double* part, AarState* S) {
  if (S->done) return;
  __shared__ double red[AB];
  const int nd = m * (m + 1) / 2 + m;
  double acc[MAXHIST * (MAXHIST + 1) / 2 + MAXHIST];
  for (int d = 0; d < nd; ++d) acc[d] = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    const double r = res[n];
    DX[(int64_t)hcol * N + n] = xp[n] - xold[n];
    DF[(int64_t)hcol * N + n] = r - fold[n];
    double df[MAXHIST];
    for (int k = 0; k < m; ++k) df[k] = DF[(int64_t)k * N + n];
    int d = 0;
    for (int a = 0; a < m; ++a)
      for (int b = a; b < m; ++b) acc[d++] += df[a] * df[b];
    for (int a = 0; a < m; ++a) acc[d++] += df[a] * r;
  }
  for (int d = 0; d < nd; ++d) {
    const double s = block_sum(acc[d], red);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nd + d] = s;
  }
  if (last_block(S)) {
    __shared__ double tot[MAXHIST * (MAXHIST + 1) / 2 + MAXHIST];
    for (int d = threadIdx.x; d < nd; d += AB) {
      double t = 0;
      for (int b = 0; b < (int)gridDim.x; ++b) t += part[(int64_t)b * nd + d];
      tot[d] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double A[MAXHIST * MAXHIST], bb[MAXHIST];
      int d = 0;
      for (int a = 0; a < m; ++a)
        for (int b = a; b < m; ++b) {
          A[a * m + b] = tot[d];
          A[b * m + a] = tot[d];
          ++d;
        }
      for (int a = 0; a < m; ++a) bb[a] = tot[d++];
      pinv_solve(m, A, bb, S->gamma);
      S->counter = 0;
    }
  }
}

// x_new = x_prev + beta*res - (DX + beta*DF) gamma
__global__ void __launch_bounds__(AB) anderson_update_kernel(int64_t N, const double* __restrict__ xp,
                                                             const double* __restrict__ res, double* __restrict__ xnew,
                                                             const double* __restrict__ DX,
                                                             const double* __restrict__ DF, int m, double beta,
                                                             const AarState* S) {
  if (S->done) return;
  double g[MAXHIST];
  for (int k = 0; k < m; ++k) g[k] = S->gamma[k];
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) {
    double corr = 0;
    for (int k = 0; k < m; ++k) corr += (DX[(int64_t)k * N + n] + beta * DF[(int64_t)k * N + n]) * g[k];
    xnew[n] = xp[n] + beta * res[n] - corr;
  }
}

__global__ void fill_kernel(double* x, int64_t N, double v) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) x[n] = v;
}

__global__ void rhs_kernel(const double* rho, const double* b, double* f, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) f[n] = -4.0 * M_PI * (rho[n] + b[n]);
}

// ---------------------------------------------------------------- host
int poisson_rhs(cyx_poisson_s* ps, const double* rho, const double* b, double* f) {
  const int64_t N = ps->N();
  hipLaunchKernelGGL(rhs_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, ps->ctx->stream, rho, b, f, N);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int poisson_run(cyx_poisson_s* ps, const double* f, double* x, int use_guess, double tol, int max_iter,
                double omega, double beta, int m, int p, cyx_aar_info* info) {
  CYX_CHECK(m >= 1 && m <= MAXHIST, "AAR history m must be in [1, %d]", MAXHIST);
  CYX_CHECK(p >= 1 && max_iter >= 1 && tol > 0, "invalid AAR parameters");
  const int64_t N = ps->N();
  hipStream_t st = ps->ctx->stream;
  const int nblk = (int)std::min<int64_t>(ceil_div(N, AB), 2048);
  const int nd = m * (m + 1) / 2 + m;
  // scratch: xalt | res0 | res1 | DX | DF | partials | state
  const size_t vec = (size_t)N * sizeof(double);
  const size_t need = vec * (3 + 2 * (size_t)m) + (size_t)nblk * nd * sizeof(double) + sizeof(AarState) + 256;
  if (ps->scratch_bytes < need) {
    if (ps->scratch) (void)hipFree(ps->scratch);
    ps->scratch = nullptr;
    ps->scratch_bytes = 0;
    if (hipMalloc((void**)&ps->scratch, need) != hipSuccess) {
      set_error("poisson scratch allocation of %zu bytes failed", need);
      return CYX_ERR_NOMEM;
    }
    ps->scratch_bytes = need;
  }
  double* xalt = ps->scratch;
  double* R[2] = {xalt + N, xalt + 2 * N};
  double* DX = xalt + 3 * N;
  double* DF = DX + (int64_t)m * N;
  double* part = DF + (int64_t)m * N;
  AarState* S = (AarState*)(((uintptr_t)(part + (int64_t)nblk * nd) + 255) & ~(uintptr_t)255);
  // X[1] = caller's x (guess), X[0] = xalt
  double* X[2] = {xalt, x};
  if (!use_guess) hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, x, N, 1.0);
  CYX_HIP(hipMemsetAsync(DX, 0, vec * 2 * m, st));
  CYX_HIP(hipMemsetAsync(S, 0, sizeof(AarState), st));
  hipLaunchKernelGGL(norm_kernel, dim3(nblk), dim3(AB), 0, st, f, N, part, S);
  PoissonOp P{};
  P.nr = ps->g.nr;
  P.nt = ps->g.nt;
  P.nz = ps->g.nz;
  P.coef = ps->coef;
  {
    double w2[FDN + 1], w1[FDN + 1];
    // same closed form as the host set-up (src/initialization.m:403-419)
    double fact[2 * FDN + 2];
    fact[0] = 1;
    for (int k = 1; k < 2 * FDN + 2; ++k) fact[k] = fact[k - 1] * k;
    w2[0] = 0;
    w1[0] = 0;
    for (int k = 1; k <= FDN; ++k) {
      const double c = fact[FDN] * fact[FDN] / (fact[FDN - k] * fact[FDN + k]);
      w2[k] = 2 * ((k % 2) ? 1.0 : -1.0) * c / (k * k);
      w1[k] = ((k % 2) ? 1.0 : -1.0) * c / k;
      w2[0] -= 2.0 / (k * k);
    }
    const double tau = ps->g.twist;
    for (int k = 0; k <= FDN; ++k) {
      P.wt2[k] = w2[k];
      P.wz2[k] = w2[k] / (ps->g.dz * ps->g.dz);
      for (int q = 0; q <= FDN; ++q) P.cx[k][q] = (-2.0 * tau * w1[k] / ps->g.dt) * (w1[q] / ps->g.dz);
    }
    P.helical = (ps->g.cell_typ == 4 || ps->g.cell_typ == 5) ? 1 : 0;
  }
  AarState hs{};
  PhaseTimer tm(ps->ctx, PH_AAR);
  int count = 1;
  const int chunk = 16;
  while (count <= max_iter) {
    const int cend = std::min(max_iter, count + chunk - 1);
    for (; count <= cend; ++count) {
      const int xb = count & 1;
      double* xp = X[xb];
      double* xo = X[xb ^ 1];  // Xold on entry, x_new on exit
      double* res = R[xb];
      double* fold = R[xb ^ 1];
      hipLaunchKernelGGL(residual_kernel, dim3(nblk), dim3(AB), 0, st, P, f, (const double*)xp, res, part, S, count,
                         max_iter, tol, xb);
      const int hcol = count > 1 ? (count - 2) % m : -1;
      if (count % p != 0) {
        hipLaunchKernelGGL(richardson_kernel, dim3(nblk), dim3(AB), 0, st, N, (const double*)xp, (const double*)res,
                           (const double*)fold, xo, DX, DF, hcol, omega, (const AarState*)S);
      } else {
        hipLaunchKernelGGL(anderson_dots_kernel, dim3(nblk), dim3(AB), 0, st, N, (const double*)xp,
                           (const double*)res, (const double*)fold, (const double*)xo, DX, DF, hcol, m, part, S);
        hipLaunchKernelGGL(anderson_update_kernel, dim3(nblk), dim3(AB), 0, st, N, (const double*)xp,
                           (const double*)res, xo, (const double*)DX, (const double*)DF, m, beta,
                           (const AarState*)S);
      }
    }
    CYX_HIP(hipGetLastError());
    CYX_HIP(hipMemcpyAsync(&hs, S, sizeof(AarState), hipMemcpyDeviceToHost, st));
    CYX_HIP(hipStreamSynchronize(st));
    if (hs.done) break;
  }
  if (hs.final_buf == 0) CYX_HIP(hipMemcpyAsync(x, xalt, vec, hipMemcpyDeviceToDevice, st));
  CYX_HIP(hipStreamSynchronize(st));
  if (info) {
    info->iterations = hs.iters;
    info->relres = hs.relres;
    info->converged = hs.done == 1;
  }
  return hs.done == 1 ? CYX_OK : CYX_ERR_NOCONV;
}

}  // namespace cyx
