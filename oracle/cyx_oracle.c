/*
 * cyx_oracle.c -- plain-C FP64 restatement of the Cyclix hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a checker and the CPU baseline of bench.py.  It
 * is never linked into or called by the product library.
 *
 * Restates (paths relative to the reference root Cyclix_MS_200220/):
 *   - the Cyclix Laplacian in its direct form with Bloch phases and radial
 *     Dirichlet truncation: Cychel/pseudochargeDensity_atom_cychel.m:4-33
 *     (stencil coefficients) over src/blochLaplacian_1d.m:16-77 (phases) and
 *     Cychel/lapIndicesValues_1d_cychel.m:17-261 (truncation / wrap);
 *   - h_nonlocal_vector_mult, src/h_nonlocal_vector_mult.m:21-43;
 *   - chebyshev_filter, src/chebyshev_filter.m:24-42;
 *   - aar with a Jacobi preconditioner, src/aar.m:33-66 and
 *     extras/AAR_old.m:45-46,73-75, on S.Lap_std (src/initialization.m:734-741).
 *
 * Complex data is interleaved (re, im) as C99 double complex; blocks are
 * column-major N x ncol.  OpenMP parallelises over z planes.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXFDN 8

typedef struct {
  int nr, nt, nz, fdn, cell_typ;
  double xin, dr, dt, dz, twist;
} cyxo_grid;

static void fd_weights(int fdn, double* w2, double* w1) {
  /* src/initialization.m:403-419 */
  double f[2 * MAXFDN + 2];
  f[0] = 1;
  for (int k = 1; k < 2 * fdn + 2; ++k) f[k] = f[k - 1] * k;
  w2[0] = 0;
  w1[0] = 0;
  for (int k = 1; k <= fdn; ++k) {
    double c = f[fdn] * f[fdn] / (f[fdn - k] * f[fdn + k]);
    w2[k] = 2 * ((k & 1) ? 1.0 : -1.0) * c / (k * k);
    w1[k] = ((k & 1) ? 1.0 : -1.0) * c / k;
    w2[0] -= 2.0 / (k * k);
  }
}

/* value of the (Bloch-extended) grid function at (i, j, k) with j, k possibly
 * outside one period; 0 outside [0, nr) radially (Dirichlet truncation). */
static inline double complex at(const cyxo_grid* g, const double complex* x, int i, int j, int k,
                                double complex ph_t, double complex ph_z) {
  if (i < 0 || i >= g->nr) return 0;
  double complex f = 1;
  while (j < 0) { j += g->nt; f *= conj(ph_t); }
  while (j >= g->nt) { j -= g->nt; f *= ph_t; }
  while (k < 0) { k += g->nz; f *= conj(ph_z); }
  while (k >= g->nz) { k -= g->nz; f *= ph_z; }
  return f * x[(long)k * g->nr * g->nt + (long)j * g->nr + i];
}

/* Lap x for one column (complex), direct stencil with separable cross term */
static void lap_column(const cyxo_grid* g, const double kpt[3], const double complex* x, double complex* out,
                       double complex* tmp) {
  const int n0 = g->fdn;
  double w2[MAXFDN + 1], w1[MAXFDN + 1];
  fd_weights(n0, w2, w1);
  const int hel = g->cell_typ == 4 || g->cell_typ == 5;
  const double tau = hel ? g->twist : 0.0;
  const double L2 = g->nt * g->dt, L3 = g->nz * g->dz;
  const double complex ph_t = kpt[1] == 0 ? 1.0 : cexp(I * kpt[1] * L2);
  const double complex ph_z = kpt[2] == 0 ? 1.0 : cexp(I * kpt[2] * L3);
  const double c1 = 1 / (g->dr * g->dr), c2 = 1 / (g->dt * g->dt), c3 = 1 / (g->dz * g->dz);
  const double c4 = tau * tau / (g->dt * g->dt), c5 = -(2 * tau) / (g->dt * g->dz), c6 = 1 / g->dr;
  const long nrt = (long)g->nr * g->nt;
  /* tmp = sum_p w1_p (x(j+p) - x(j-p)) : theta first derivative (unscaled) */
  if (hel) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < g->nz; ++k)
      for (int j = 0; j < g->nt; ++j)
        for (int i = 0; i < g->nr; ++i) {
          double complex s = 0;
          for (int p = 1; p <= n0; ++p) s += w1[p] * (at(g, x, i, j + p, k, ph_t, ph_z) - at(g, x, i, j - p, k, ph_t, ph_z));
          tmp[k * nrt + (long)j * g->nr + i] = s;
        }
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->nt; ++j)
      for (int i = 0; i < g->nr; ++i) {
        const double rr = g->xin + i * g->dr;
        double complex b = w2[0] * (c1 + c2 / (rr * rr) + c3 + c4) * at(g, x, i, j, k, ph_t, ph_z);
        for (int p = 1; p <= n0; ++p) {
          b += (c1 * w2[p] + c6 * w1[p] / rr) * at(g, x, i + p, j, k, ph_t, ph_z) +
               (c1 * w2[p] - c6 * w1[p] / rr) * at(g, x, i - p, j, k, ph_t, ph_z) +
               (c2 / (rr * rr) + c4) * w2[p] * (at(g, x, i, j + p, k, ph_t, ph_z) + at(g, x, i, j - p, k, ph_t, ph_z)) +
               c3 * w2[p] * (at(g, x, i, j, k + p, ph_t, ph_z) + at(g, x, i, j, k - p, ph_t, ph_z));
          if (hel) /* sum_q c5 w1_p w1_q [...] = c5 w1_q (T(k+q) - T(k-q)) with T the theta derivative */
            b += c5 * w1[p] * (at(g, tmp, i, j, k + p, ph_t, ph_z) - at(g, tmp, i, j, k - p, ph_t, ph_z));
        }
        out[k * nrt + (long)j * g->nr + i] = b;
      }
}

typedef struct {
  int n_atom, n_image;
  const int* angnum;
  const double* gamma;
  const int* image_atom;
  const int* image_nnodes;
  const double* image_disp;
  const int* rc_pos;
  const double* chi;
} cyxo_proj;

/* H x for one column, src/h_nonlocal_vector_mult.m:21-43 */
static void h_column(const cyxo_grid* g, const double kpt[3], const double* veff, const double* W, const cyxo_proj* P,
                     const double complex* x, double complex* hx, double complex* tmp) {
  const long N = (long)g->nr * g->nt * g->nz;
  lap_column(g, kpt, x, hx, tmp);
#pragma omp parallel for schedule(static)
  for (long n = 0; n < N; ++n) hx[n] = -0.5 * hx[n] + veff[n] * x[n];
  if (!P || P->n_atom == 0) return;
  const int any_k = kpt[0] != 0 || kpt[1] != 0 || kpt[2] != 0;
  long off = 0, choff = 0;
  /* image offsets per atom are contiguous in the flattened description */
  long* ioff = malloc(sizeof(long) * (P->n_image + 1));
  long* coff = malloc(sizeof(long) * (P->n_image + 1));
  for (int im = 0; im < P->n_image; ++im) {
    ioff[im] = off;
    coff[im] = choff;
    off += P->image_nnodes[im];
    choff += (long)P->image_nnodes[im] * P->angnum[P->image_atom[im]];
  }
  int pb = 0;
  for (int J = 0; J < P->n_atom; ++J) {
    const int na = P->angnum[J];
    double complex C[16];
    for (int l = 0; l < na; ++l) C[l] = 0;
    for (int im = 0; im < P->n_image; ++im) {
      if (P->image_atom[im] != J) continue;
      const double* d = P->image_disp + 3 * im;
      const double complex ph = any_k ? cexp(I * (kpt[0] * d[0] + kpt[1] * d[1] + kpt[2] * d[2])) : 1.0;
      for (int l = 0; l < na; ++l) {
        double complex s = 0;
        for (int e = 0; e < P->image_nnodes[im]; ++e) {
          const int node = P->rc_pos[ioff[im] + e];
          s += P->chi[coff[im] + (long)l * P->image_nnodes[im] + e] * W[node] * x[node];
        }
        C[l] += s * ph;
      }
    }
    for (int l = 0; l < na; ++l) C[l] *= P->gamma[pb + l];
    for (int im = 0; im < P->n_image; ++im) {
      if (P->image_atom[im] != J) continue;
      const double* d = P->image_disp + 3 * im;
      const double complex ph = any_k ? cexp(-I * (kpt[0] * d[0] + kpt[1] * d[1] + kpt[2] * d[2])) : 1.0;
      for (int e = 0; e < P->image_nnodes[im]; ++e) {
        const int node = P->rc_pos[ioff[im] + e];
        double complex s = 0;
        for (int l = 0; l < na; ++l) s += P->chi[coff[im] + (long)l * P->image_nnodes[im] + e] * C[l];
        hx[node] += s * ph;
      }
    }
    pb += na;
  }
  free(ioff);
  free(coff);
}

int cyxo_h_apply(const cyxo_grid* g, const double kpt[3], const double* veff, const double* W, const cyxo_proj* P,
                 const double complex* X, double complex* HX, int ncol) {
  const long N = (long)g->nr * g->nt * g->nz;
  double complex* tmp = malloc(sizeof(double complex) * N);
  if (!tmp) return -4;
  for (int s = 0; s < ncol; ++s) h_column(g, kpt, veff, W, P, X + s * N, HX + s * N, tmp);
  free(tmp);
  return 0;
}

/* src/chebyshev_filter.m:24-42; X is overwritten, result in Y */
int cyxo_chebyshev_filter(const cyxo_grid* g, const double kpt[3], const double* veff, const double* W,
                          const cyxo_proj* P, double complex* X, double complex* Y, int ncol, int m, double a,
                          double b, double a0) {
  const long N = (long)g->nr * g->nt * g->nz;
  double complex* tmp = malloc(sizeof(double complex) * N);
  double complex* hx = malloc(sizeof(double complex) * N);
  if (!tmp || !hx) return -4;
  const double e = (b - a) / 2, c = (b + a) / 2;
  for (int s = 0; s < ncol; ++s) {
    double complex* x = X + s * N;
    double complex* y = Y + s * N;
    double sigma = e / (a0 - c);
    const double sigma1 = sigma, gamma = 2 / sigma1;
    h_column(g, kpt, veff, W, P, x, hx, tmp);
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) y[n] = (sigma1 / e) * (hx[n] - c * x[n]);
    for (int it = 2; it <= m; ++it) {
      const double sigma2 = 1 / (gamma - sigma);
      h_column(g, kpt, veff, W, P, y, hx, tmp);
#pragma omp parallel for schedule(static)
      for (long n = 0; n < N; ++n) {
        const double complex ynew = (2 * sigma2 / e) * (hx[n] - c * y[n]) - (sigma * sigma2) * x[n];
        x[n] = y[n];
        y[n] = ynew;
      }
      sigma = sigma2;
    }
  }
  free(tmp);
  free(hx);
  return 0;
}

/* ---------------------------------------------------------------- AAR */
static void lap_real(const cyxo_grid* g, const double* x, double* out, double complex* cx, double complex* co,
                     double complex* tmp) {
  const long N = (long)g->nr * g->nt * g->nz;
  const double k0[3] = {0, 0, 0};
  for (long n = 0; n < N; ++n) cx[n] = x[n];
  lap_column(g, k0, cx, co, tmp);
  for (long n = 0; n < N; ++n) out[n] = creal(co[n]);
}

/* symmetric pseudo-inverse solve by Jacobi eigen-decomposition (MATLAB pinv
 * tolerance max(size)*norm*eps) */
static void pinv_solve(int m, double* A, const double* bb, double* out) {
  double V[16 * 16];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) V[i * m + j] = i == j;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < m * m; ++i) tot += A[i] * A[i];
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        if (i != j) off += A[i * m + j] * A[i * m + j];
    if (off <= 1e-32 * tot || off == 0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[p * m + q];
        if (apq == 0) continue;
        const double th = (A[q * m + q] - A[p * m + p]) / (2 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
        const double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double a1 = A[k * m + p], a2 = A[k * m + q];
          A[k * m + p] = c * a1 - s * a2;
          A[k * m + q] = s * a1 + c * a2;
        }
        for (int k = 0; k < m; ++k) {
          const double a1 = A[p * m + k], a2 = A[q * m + k];
          A[p * m + k] = c * a1 - s * a2;
          A[q * m + k] = s * a1 + c * a2;
        }
        for (int k = 0; k < m; ++k) {
          const double v1 = V[k * m + p], v2 = V[k * m + q];
          V[k * m + p] = c * v1 - s * v2;
          V[k * m + q] = s * v1 + c * v2;
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
    for (int k = 0; k < m; ++k) vb += V[k * m + e] * bb[k];
    vb /= lam;
    for (int k = 0; k < m; ++k) out[k] += V[k * m + e] * vb;
  }
}

/* x = aar(Lap_std, f, x0, tol, max_iter, omega, beta, m, p) with Jacobi
 * preconditioning; returns iterations, writes relres. src/aar.m:33-66 */
int cyxo_aar_jacobi(const cyxo_grid* g, const double* f, double* x_io, int use_guess, double tol, int max_iter,
                    double omega, double beta, int m, int p, double* relres_out) {
  const long N = (long)g->nr * g->nt * g->nz;
  if (m > 16) return -1;
  double w2[MAXFDN + 1], w1[MAXFDN + 1];
  fd_weights(g->fdn, w2, w1);
  const double tau = (g->cell_typ == 4 || g->cell_typ == 5) ? g->twist : 0.0;
  double* diag = malloc(sizeof(double) * g->nr);
  for (int i = 0; i < g->nr; ++i) {
    const double r = g->xin + i * g->dr;
    diag[i] = w2[0] / (g->dr * g->dr) + w2[0] * (tau * tau + 1 / (r * r)) / (g->dt * g->dt) + w2[0] / (g->dz * g->dz);
  }
  double *DX = calloc((size_t)N * m, sizeof(double)), *DF = calloc((size_t)N * m, sizeof(double));
  double *xp = malloc(sizeof(double) * N), *xold = malloc(sizeof(double) * N), *fold = malloc(sizeof(double) * N);
  double *res = malloc(sizeof(double) * N), *ax = malloc(sizeof(double) * N), *xret = malloc(sizeof(double) * N);
  double complex *cx = malloc(sizeof(double complex) * N), *co = malloc(sizeof(double complex) * N),
                 *tmp = malloc(sizeof(double complex) * N);
  double nb = 0;
  for (long n = 0; n < N; ++n) nb += f[n] * f[n];
  nb = 1 / sqrt(nb);
  for (long n = 0; n < N; ++n) xp[n] = use_guess ? x_io[n] : 1.0;
  double relres = 1 + tol;
  int count = 1;
  while (count <= max_iter && relres > tol) {
    lap_real(g, xp, ax, cx, co, tmp);
    double s = 0;
    for (long n = 0; n < N; ++n) {
      const double r1 = f[n] - ax[n];
      s += r1 * r1;
      res[n] = r1 / diag[n % g->nr];
    }
    relres = sqrt(s) * nb;
    if (count > 1) {
      const int k = (count - 2) % m;
      for (long n = 0; n < N; ++n) {
        DX[k * N + n] = xp[n] - xold[n];
        DF[k * N + n] = res[n] - fold[n];
      }
    }
    memcpy(xold, xp, sizeof(double) * N);
    memcpy(fold, res, sizeof(double) * N);
    memcpy(xret, xp, sizeof(double) * N);
    if (count % p != 0) {
      for (long n = 0; n < N; ++n) xp[n] += omega * res[n];
    } else {
      double A[16 * 16], bb[16], gm[16];
      for (int a = 0; a < m; ++a) {
        double t = 0;
        for (long n = 0; n < N; ++n) t += DF[a * N + n] * res[n];
        bb[a] = t;
        for (int c = a; c < m; ++c) {
          double u = 0;
          for (long n = 0; n < N; ++n) u += DF[a * N + n] * DF[c * N + n];
          A[a * m + c] = A[c * m + a] = u;
        }
      }
      pinv_solve(m, A, bb, gm);
      for (long n = 0; n < N; ++n) {
        double corr = 0;
        for (int k = 0; k < m; ++k) corr += (DX[k * N + n] + beta * DF[k * N + n]) * gm[k];
        xp[n] = xp[n] + beta * res[n] - corr;
      }
    }
    ++count;
  }
  memcpy(x_io, xret, sizeof(double) * N);
  *relres_out = relres;
  free(diag); free(DX); free(DF); free(xp); free(xold); free(fold); free(res); free(ax); free(xret);
  free(cx); free(co); free(tmp);
  return count - 1;
}
