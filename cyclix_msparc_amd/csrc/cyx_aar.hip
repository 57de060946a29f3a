// Jacobi-preconditioned Alternating Anderson-Richardson Poisson solve
// (src/aar.m:33-72, src/poissonSolve.m:17-51; Jacobi preconditioner of
// extras/AAR_old.m:45-46,73-75).  Operator: S.Lap_std at k = 0
// (src/initialization.m:734-741), applied matrix-free.
//
// The 2.2 MB-per-vector problem is latency-bound, so an iteration is kept to
// as few dependent launches as possible:
//   Richardson step (one launch): the step kernel forms T = (-2 tau D_theta) x
//   (the helical cross term's theta half) for the planes it stages, then
//   res1 = f - Lap x, res = res1 ./ diag, the history columns,
//   x_new = x + omega res (written speculatively into another x buffer) and
//   per-block partials of ||res1||^2; one extra block finalises the previous
//   Richardson iteration (fixed-order sum of its partials, relres, the
//   convergence flag).  A step block takes 8 z planes of one 64-point plane
//   segment and stages x (with its z and r halo), T and the radial
//   coefficients in LDS, so only the theta taps are read through the vector
//   cache.
//   Anderson step (three launches): the step kernel with the DF'DF and DF'res
//   partials, a one-block finalize that also computes
//   gamma = pinv(DF'DF) DF'res on one wavefront, and
//   x_new = x + beta res - (DX + beta DF) gamma.
// theta_kernel (the standalone T) serves the learned correction and carries
// a finalise just before a poll.
// Partials are stored column-major, part[d][block].  Step kernels of an
// iteration after convergence store nothing, so the host enqueues iterations
// in chunks and polls a device flag between chunks.  Every reduction has a
// fixed order: a solve is bitwise reproducible.
#include <cmath>

#include "cyx_internal.hpp"
#include "cyx_reduce.hpp"

namespace cyx {

struct AarState {
  int done;       // 1 converged, 2 hit max_iter
  int iters;      // count at exit (src/aar.m:68-72 reports count-1 after the loop)
  int final_buf;  // which x buffer holds the returned iterate
  int pinv_eig;   // Anderson steps that took the eigen-decomposition pseudo-inverse
  double relres;  // relres of the returned iterate
  double nb_inv;  // 1 / ||f||
  double gamma[16];
};

struct PoissonOp {
  int nr, nt, nz, N;
  int helical;
  const double* coefT;  // [RCOEF][nr]: c+[6], c-[6], c_theta, diag (Lap_std, no -1/2), radius fastest;
                        // radial taps outside [0, nr) are zero
  double w2[FDN + 1];   // w2_p (theta second derivative, times c_theta per radius)
  double wz2[FDN + 1];  // w2_q / dz^2
  double wt1[FDN + 1];  // -2 tau w1_p / dt
  double wz1[FDN + 1];  // w1_q / dz
};

__device__ inline int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Per-point kernels: 1-D blocks of PB threads, one grid point per thread in
// the r-fastest linear order (N < 2^31 is checked on the host).
constexpr int PB = 512;

struct Pt {
  int i, j, k, n;
  bool ok;
};
__device__ inline Pt point_of(const PoissonOp& P) {
  Pt q;
  q.n = blockIdx.x * PB + threadIdx.x;
  q.ok = q.n < P.N;
  const int nrt = P.nr * P.nt;
  q.k = q.n / nrt;
  const int rem = q.n - q.k * nrt;
  q.j = rem / P.nr;
  q.i = rem - q.j * P.nr;
  return q;
}

__device__ inline double diag_at(const PoissonOp& P, int i) { return P.coefT[(2 * FDN + 1) * P.nr + i]; }

// Finalise of a Richardson iteration (src/aar.m:43-45): relres of x_prev
// from the nblk partials of ||res1||^2, convergence and iteration-cap flags.
struct FinArgs {
  const double* part;
  int nblk, count, max_iter, xbuf;
  double tol;
  AarState* S;
};

__device__ void finalize_richardson(const FinArgs& F) {
  if (F.S->done) return;
  __shared__ double red[PB / 64];
  const double tot = column_sums_all<1, PB>(F.part, F.nblk, red);
  if (threadIdx.x == 0) {
    const double relres = sqrt(tot) * F.S->nb_inv;
    F.S->relres = relres;
    F.S->iters = F.count;
    F.S->final_buf = F.xbuf;
    if (relres <= F.tol)
      F.S->done = 1;
    else if (F.count >= F.max_iter)
      F.S->done = 2;
  }
}

// Blocks [0, ntb): T[n] = sum_p wt1_p (x(j+p) - x(j-p)).  Block ntb, when
// launched, finalises the previous Richardson iteration.  T is scratch, so
// the theta blocks run whatever the convergence state.
__global__ void __launch_bounds__(PB) theta_kernel(PoissonOp P, const double* __restrict__ x, double* __restrict__ T,
                                                   int ntb, int pair, FinArgs F) {
  if ((int)blockIdx.x == ntb) {
    finalize_richardson(F);
    return;
  }
  if (pair) {
    // two consecutive radii per thread, 16-byte loads and stores
    const int hr = P.nr >> 1;
    const int pi = blockIdx.x * PB + threadIdx.x;
    const int row = pi / hr;  // k * nt + j
    if (row >= P.nz * P.nt) return;
    const int i = 2 * (pi - row * hr), j = row % P.nt;
    const int n = row * P.nr + i, base = n - j * P.nr;
    double ta = 0, tb = 0;
#pragma unroll
    for (int p = 1; p <= FDN; ++p) {
      const double2 u = *reinterpret_cast<const double2*>(x + base + wrapi(j + p, P.nt) * P.nr);
      const double2 d = *reinterpret_cast<const double2*>(x + base + wrapi(j - p, P.nt) * P.nr);
      ta = __builtin_fma(P.wt1[p], u.x - d.x, ta);
      tb = __builtin_fma(P.wt1[p], u.y - d.y, tb);
    }
    *reinterpret_cast<double2*>(T + n) = make_double2(ta, tb);
    return;
  }
  const Pt q = point_of(P);
  if (!q.ok) return;
  const int base = q.n - q.j * P.nr;  // (i, 0, k)
  double t = 0;
#pragma unroll
  for (int p = 1; p <= FDN; ++p)
    t = __builtin_fma(P.wt1[p], x[base + wrapi(q.j + p, P.nt) * P.nr] - x[base + wrapi(q.j - p, P.nt) * P.nr], t);
  T[q.n] = t;
}

// Launch shape of theta_kernel: two radii per thread when nr is even and both
// vectors are 16-byte aligned
struct ThetaShape {
  int ntb, pair;
};
static ThetaShape theta_shape(const PoissonOp& P, const double* x, const double* T) {
  const int pair = ((P.nr & 1) == 0 && (((uintptr_t)x | (uintptr_t)T) & 15) == 0) ? 1 : 0;
  return ThetaShape{(int)ceil_div((int64_t)P.N / (pair ? 2 : 1), PB), pair};
}

__device__ inline double lap_at(const PoissonOp& P, const double* __restrict__ x, const double* __restrict__ T,
                                const Pt& q) {
  const int nr = P.nr, nrt = P.nr * P.nt, i = q.i, j = q.j, k = q.k, n = q.n;
  const double* cf = P.coefT + i;
  double acc = cf[(2 * FDN + 1) * nr] * x[n];
#pragma unroll
  for (int p = 1; p <= FDN; ++p) {
    if (i + p < nr) acc = __builtin_fma(cf[(p - 1) * nr], x[n + p], acc);
    if (i - p >= 0) acc = __builtin_fma(cf[(FDN + p - 1) * nr], x[n - p], acc);
  }
  const int rowk = k * nrt + i;
  double st = 0.0;
#pragma unroll
  for (int p = 1; p <= FDN; ++p)
    st = __builtin_fma(P.w2[p], x[rowk + wrapi(j + p, P.nt) * nr] + x[rowk + wrapi(j - p, P.nt) * nr], st);
  acc = __builtin_fma(cf[2 * FDN * nr], st, acc);
  const int colj = j * nr + i;
#pragma unroll
  for (int qq = 1; qq <= FDN; ++qq) {
    const int kp = wrapi(k + qq, P.nz) * nrt + colj;
    const int km = wrapi(k - qq, P.nz) * nrt + colj;
    acc = __builtin_fma(P.wz2[qq], x[kp] + x[km], acc);
    if (P.helical) acc = __builtin_fma(P.wz1[qq], T[kp] - T[km], acc);
  }
  return acc;
}

// ---- learned correction with exact line search (Anderson iterations):
//   r1 = f - Lap x,  d = r1 ./ diag + U(r1 ./ diag),  w = Lap d,
//   delta = <w, r1> / <w, w>,  x += delta d
// delta minimises ||f - Lap (x + delta d)||, so the step never raises the
// residual norm whatever the network returns.

// r1 = f - Lap x
__global__ void __launch_bounds__(PB) resid_kernel(PoissonOp P, const double* __restrict__ f,
                                                   const double* __restrict__ x, const double* __restrict__ T,
                                                   double* __restrict__ r1, const AarState* S) {
  if (S->done) return;
  const Pt q = point_of(P);
  if (!q.ok) return;
  r1[q.n] = f[q.n] - lap_at(P, x, T, q);
}

__global__ void __launch_bounds__(PB) ls_dir_kernel(PoissonOp P, const double* __restrict__ r1,
                                                    const float* __restrict__ u, double* __restrict__ d,
                                                    const AarState* S) {
  if (S->done) return;
  const Pt q = point_of(P);
  if (!q.ok) return;
  d[q.n] = r1[q.n] / diag_at(P, q.i) + (double)u[q.n];
}

__global__ void __launch_bounds__(PB) ls_apply_kernel(PoissonOp P, const double* __restrict__ d,
                                                      const double* __restrict__ T, const double* __restrict__ r1,
                                                      double* __restrict__ part, const AarState* S) {
  if (S->done) return;
  __shared__ double red[(PB / 64) * 2];
  const Pt q = point_of(P);
  double v[2] = {0, 0};
  if (q.ok) {
    const double w = lap_at(P, d, T, q);
    v[0] = w * r1[q.n];
    v[1] = w * w;
  }
  const double s = block_sums<2, PB>(v, red);
  if (threadIdx.x < 2) part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(AB) ls_finalize_kernel(const double* part, int nblk, AarState* S) {
  if (S->done) return;
  __shared__ double tot[2];
  column_sums(part, nblk, 2, tot);
  __syncthreads();
  if (threadIdx.x == 0) S->gamma[15] = tot[1] > 0 ? tot[0] / tot[1] : 0.0;
}

__global__ void __launch_bounds__(PB) ls_axpy_kernel(PoissonOp P, double* __restrict__ x,
                                                     const double* __restrict__ d, const AarState* S) {
  if (S->done) return;
  const Pt q = point_of(P);
  if (!q.ok) return;
  x[q.n] += S->gamma[15] * d[q.n];
}

__global__ void __launch_bounds__(AB) norm_kernel(const double* __restrict__ f, int64_t N, double* part) {
  __shared__ double red[AB];
  double a = 0;
  for (int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x; n < N; n += (int64_t)gridDim.x * AB) a += f[n] * f[n];
  const double s = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(AB) norm_finalize_kernel(const double* part, int nblk, AarState* S) {
  __shared__ double tot[1];
  column_sums(part, nblk, 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) S->nb_inv = 1.0 / sqrt(tot[0]);
}

struct StepArgs {
  const double* f;
  const double* xp;     // x_prev (its residual is measured)
  const double* xold;   // the iterate before x_prev (history column)
  double* xnew;         // x_new of a Richardson step
  double* res;          // preconditioned residual of x_prev (out)
  const double* fold;   // previous preconditioned residual
  double* DX;
  double* DF;
  double* part;         // [nd][nstep]
  int nstep;            // step blocks (the partials' stride)
  AarState* S;
  int hcol;             // history column written this iteration (-1: none)
  int anderson;         // 1 on Anderson steps
  int count, max_iter, xbuf;
  double tol, omega;
};

// Step kernel.  MH: compile-time history size of the Anderson dot products
// (0: Richardson step, which writes x_new; MH > 0: Anderson step, which
// writes the partials of DF'DF and DF'res instead).
// Brick form: block b covers plane segment seg (64
// consecutive points of a plane in the r-fastest order) of planes
// 8 zc .. 8 zc + 7, wave w one plane.  LDS holds x at planes k - FDN ..
// k + FDN of the block's planes over the segment and an r halo of FDN on
// each side, the helical theta derivative T = (-2 tau D_theta) x at the same
// planes over the segment and the theta second-derivative sum of the block's
// own planes (both formed here from the same theta-neighbour loads), and the
// 14 coefficients of the segment's radii.  The per-point vectors (f, the
// previous iterate and residual, the Anderson history) are loaded before the
// staging barrier, so their latency overlaps it.  Halo positions outside the plane are clamped into
// it: they only meet radial taps whose coefficient is zero.  With fin set,
// one block more (the last) finalises the previous Richardson iteration; the
// iterates are triple-buffered, so this step's stores never touch the
// iterate that finalise may select.
constexpr int BZ = PB / 64;           // planes per step block
constexpr int BSEG = 64;              // points per plane segment
constexpr int BZH = BZ + 2 * FDN;     // staged planes
constexpr int BXW = BSEG + 2 * FDN;   // staged x row with r halo

struct BrickLds {
  double x[BZH][BXW];
  double t[BZH][BSEG];
  double s[BZ][BSEG];  // sum_p w2_p (x(j+p) + x(j-p)) of the block's planes
  double c[RCOEF][BSEG];
};

// Richardson and short-history Anderson steps are held to 80 VGPRs, so that
// three blocks share a CU and a production grid (~540 blocks) is resident at
// once; longer histories need more registers for their block sums
template <int MH>
constexpr int brick_waves_per_eu() {
  return MH < 5 ? 6 : 1;
}

template <int MH>
__global__ void __launch_bounds__(PB) __attribute__((amdgpu_waves_per_eu(brick_waves_per_eu<MH>())))
aar_step_brick_kernel(PoissonOp P, StepArgs A, FinArgs F, int fin) {
  constexpr int ND = MH * (MH + 1) / 2 + MH;
  if (fin && (int)blockIdx.x == A.nstep) {
    finalize_richardson(F);
    return;
  }
  __shared__ double red[(PB / 64) * (1 + ND)];
  __shared__ BrickLds L;
  const int done = A.S->done;  // stores are skipped after convergence; the loads need not wait for it
  const int64_t N = P.N;
  const int nrt = P.nr * P.nt;
  const int nseg = (nrt + BSEG - 1) / BSEG;
  const int seg = blockIdx.x % nseg, k0 = (blockIdx.x / nseg) * BZ;
  const int p0 = seg * BSEG;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k = k0 + w, pos = p0 + lane;
  const bool ok = k < P.nz && pos < nrt;
  const int n = k * nrt + pos;
  // per-point loads first (they do not depend on the staging)
  double fv = 0, xold = 0, rold = 0;
  double dfo[MH > 0 ? MH : 1];
  if (ok) {
    fv = A.f[n];
    if (A.hcol >= 0) {
      xold = A.xold[n];
      rold = A.fold[n];
    }
    if constexpr (MH > 0) {
#pragma unroll
      for (int a = 0; a < MH; ++a) dfo[a] = A.DF[(int64_t)a * N + n];
    }
  }
  // staging: every load is issued unconditionally (indices clamped into
  // range, only the LDS stores are predicated) and the x and coefficient
  // values are stored after the theta work, so that they travel with the
  // first trip's theta loads: a helical block waits on three memory round
  // trips (one per theta trip) instead of one per trip of every loop
  constexpr int NXL = (BZH * BXW + PB - 1) / PB;
  constexpr int NCL = (RCOEF * BSEG + PB - 1) / PB;
  constexpr int NTL = (BZH * BSEG + PB - 1) / PB;
  double xr[NXL], cr[NCL];
#pragma unroll
  for (int it = 0; it < NXL; ++it) {
    const int e = min((int)threadIdx.x + it * PB, BZH * BXW - 1);
    const int pl = e / BXW, off = e % BXW;
    const int kx = ((k0 - FDN + pl) % P.nz + P.nz) % P.nz;
    const int px = min(max(p0 - FDN + off, 0), nrt - 1);
    xr[it] = A.xp[kx * nrt + px];
  }
#pragma unroll
  for (int it = 0; it < NCL; ++it) {
    const int e = min((int)threadIdx.x + it * PB, RCOEF * BSEG - 1);
    const int cc = e / BSEG, off = e % BSEG;
    cr[it] = P.coefT[cc * P.nr + min(p0 + off, nrt - 1) % P.nr];
  }
  // theta neighbours of staged point (pl, off) (planes past nz, in a partial
  // last brick, are wrapped and never emitted)
  auto theta_loads = [&](int pl, int off, double (&xu)[FDN], double (&xd)[FDN]) {
    const int kt = ((k0 - FDN + pl) % P.nz + P.nz) % P.nz;
    const int pt = min(p0 + off, nrt - 1);
    const int jt = pt / P.nr, base = kt * nrt + (pt - jt * P.nr);
#pragma unroll
    for (int q = 1; q <= FDN; ++q) {
      xu[q - 1] = A.xp[base + wrapi(jt + q, P.nt) * P.nr];
      xd[q - 1] = A.xp[base + wrapi(jt - q, P.nt) * P.nr];
    }
  };
  auto theta_sum = [&](const double (&xu)[FDN], const double (&xd)[FDN]) {
    double st = 0;
#pragma unroll
    for (int q = 1; q <= FDN; ++q) st = __builtin_fma(P.w2[q], xu[q - 1] + xd[q - 1], st);
    return st;
  };
  if (P.helical) {
    // T at every staged plane, as theta_kernel forms it, and the theta sum
    // at the block's own planes, from one set of neighbour loads
#pragma unroll
    for (int it = 0; it < NTL; ++it) {
      const int e0 = (int)threadIdx.x + it * PB;
      const int e = min(e0, BZH * BSEG - 1);
      const int pl = e / BSEG, off = e % BSEG;
      double xu[FDN], xd[FDN];
      theta_loads(pl, off, xu, xd);
      if (e0 < BZH * BSEG) {
        double t = 0;
#pragma unroll
        for (int q = 1; q <= FDN; ++q) t = __builtin_fma(P.wt1[q], xu[q - 1] - xd[q - 1], t);
        L.t[pl][off] = t;
        if (pl >= FDN && pl < FDN + BZ) L.s[pl - FDN][off] = theta_sum(xu, xd);
      }
    }
  } else {
    // the theta sum at the block's own planes: one point per thread
    double xu[FDN], xd[FDN];
    theta_loads(FDN + w, lane, xu, xd);
    L.s[w][lane] = theta_sum(xu, xd);
  }
#pragma unroll
  for (int it = 0; it < NXL; ++it) {
    const int e = threadIdx.x + it * PB;
    if (e < BZH * BXW) L.x[e / BXW][e % BXW] = xr[it];
  }
#pragma unroll
  for (int it = 0; it < NCL; ++it) {
    const int e = threadIdx.x + it * PB;
    if (e < RCOEF * BSEG) L.c[e / BSEG][e % BSEG] = cr[it];
  }
  __syncthreads();
  double v[1 + ND];
#pragma unroll
  for (int d = 0; d < 1 + ND; ++d) v[d] = 0;
  double x = 0, r = 0;
  if (ok) {
    const int xc = FDN + lane, zc = FDN + w;
    x = L.x[zc][xc];
    double acc = L.c[2 * FDN + 1][lane] * x;
#pragma unroll
    for (int p = 1; p <= FDN; ++p) {
      acc = __builtin_fma(L.c[p - 1][lane], L.x[zc][xc + p], acc);
      acc = __builtin_fma(L.c[FDN + p - 1][lane], L.x[zc][xc - p], acc);
    }
    acc = __builtin_fma(L.c[2 * FDN][lane], L.s[w][lane], acc);
#pragma unroll
    for (int q = 1; q <= FDN; ++q) {
      acc = __builtin_fma(P.wz2[q], L.x[zc + q][xc] + L.x[zc - q][xc], acc);
      if (P.helical) acc = __builtin_fma(P.wz1[q], L.t[zc + q][lane] - L.t[zc - q][lane], acc);
    }
    const double r1 = fv - acc;
    v[0] = r1 * r1;
    r = r1 / L.c[2 * FDN + 1][lane];
  }
  if (done) return;
  if (ok) {
    A.res[n] = r;
    if (A.hcol >= 0) {
      A.DX[(int64_t)A.hcol * N + n] = x - xold;
      A.DF[(int64_t)A.hcol * N + n] = r - rold;
    }
    if constexpr (MH == 0) {
      A.xnew[n] = __builtin_fma(A.omega, r, x);
    } else {
      double df[MH];
#pragma unroll
      for (int a = 0; a < MH; ++a) df[a] = a == A.hcol ? r - rold : dfo[a];
      int d = 1;
#pragma unroll
      for (int a = 0; a < MH; ++a)
#pragma unroll
        for (int b = a; b < MH; ++b) v[d++] = df[a] * df[b];
#pragma unroll
      for (int a = 0; a < MH; ++a) v[d++] = df[a] * r;
    }
  }
  const double s = block_sums<1 + ND, PB>(v, red);
  if ((int)threadIdx.x < 1 + ND) A.part[(int64_t)threadIdx.x * A.nstep + blockIdx.x] = s;
}

// one block: fixed-order sums of an Anderson step's partials, convergence
// decision (src/aar.m:43-45) and, unless converged, gamma = pinv(DF'DF)
// DF'res (src/aar.m:59) on wavefront 0: by Cholesky when DF'DF is certified
// full rank under pinv's tolerance (then pinv = inverse), else by the
// eigen-decomposition pseudo-inverse
template <int MH>
__global__ void __launch_bounds__(AB) aar_finalize_kernel(const double* part, int nblk, int count, int max_iter,
                                                          int xbuf, double tol, AarState* S) {
  if (S->done) return;
  constexpr int ND = MH * (MH + 1) / 2 + MH;
  __shared__ double tot[ND + 1];
  __shared__ double red[(AB / 64) * (ND + 1)];
  __shared__ PinvShared<MH> W;
  __shared__ int go;
  {
    const double sd = column_sums_all<1 + ND, AB>(part, nblk, red);
    if ((int)threadIdx.x < 1 + ND) tot[threadIdx.x] = sd;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double relres = sqrt(tot[0]) * S->nb_inv;
    S->relres = relres;
    S->iters = count;
    S->final_buf = xbuf;
    if (relres <= tol)
      S->done = 1;
    else if (count >= max_iter)
      S->done = 2;
    go = S->done ? 0 : 1;
  }
  __syncthreads();
  if (!go || threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  for (int e = lane; e < MH * MH; e += 64) {
    const int a = e / MH, b = e % MH, lo = a < b ? a : b, hi = a < b ? b : a;
    // packed upper-triangle index of (lo, hi): rows lo' < lo hold MH - lo' entries
    W.A[e] = tot[1 + lo * MH - lo * (lo - 1) / 2 + (hi - lo)];
  }
  if (lane < MH) W.b[lane] = tot[1 + MH * (MH + 1) / 2 + lane];
  wave_sync();
  if (!chol_solve_regs<MH>(W, lane)) {
    pinv_solve_wave<MH>(W, lane);
    if (lane == 0) S->pinv_eig += 1;
  }
  if (lane < MH) S->gamma[lane] = W.out[lane];
}

// x_new = x_prev + beta res - (DX + beta DF) gamma   (src/aar.m:59-60)
template <int MH>
__global__ void __launch_bounds__(PB) anderson_update_kernel(PoissonOp P, const double* __restrict__ xp,
                                                             const double* __restrict__ res, double* __restrict__ xnew,
                                                             const double* __restrict__ DX,
                                                             const double* __restrict__ DF, double beta,
                                                             const AarState* S) {
  if (S->done) return;
  const Pt q = point_of(P);
  if (!q.ok) return;
  const int64_t N = P.N;
  const int n = q.n;
  double corr = 0;
#pragma unroll
  for (int k = 0; k < MH; ++k) corr += (DX[(int64_t)k * N + n] + beta * DF[(int64_t)k * N + n]) * S->gamma[k];
  xnew[n] = xp[n] + beta * res[n] - corr;
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

// step blocks (brick form) and point blocks (one point per thread)
static int step_blocks(const PoissonOp& P) {
  return ceil_div(P.nr * P.nt, BSEG) * ceil_div(P.nz, BZ);
}

template <int MH>
static void launch_anderson(int nblk, int nstep, hipStream_t st, const PoissonOp& P, const StepArgs& A, double beta,
                            double* xnew, const FinArgs& F, bool fin) {
  hipLaunchKernelGGL(aar_step_brick_kernel<MH>, dim3(nstep + (fin ? 1 : 0)), dim3(PB), 0, st, P, A, F, fin ? 1 : 0);
  hipLaunchKernelGGL(aar_finalize_kernel<MH>, dim3(1), dim3(AB), 0, st, (const double*)A.part, nstep, A.count,
                     A.max_iter, A.xbuf, A.tol, A.S);
  hipLaunchKernelGGL(anderson_update_kernel<MH>, dim3(nblk), dim3(PB), 0, st, P, A.xp, (const double*)A.res, xnew,
                     (const double*)A.DX, (const double*)A.DF, beta, (const AarState*)A.S);
}

static PoissonOp make_op(const cyx_poisson_s* ps) {
  PoissonOp P{};
  P.nr = ps->g.nr;
  P.nt = ps->g.nt;
  P.nz = ps->g.nz;
  P.N = (int)ps->N();
  P.coefT = ps->coefT;
  P.helical = (ps->g.cell_typ == 4 || ps->g.cell_typ == 5) ? 1 : 0;
  {
    double fact[2 * FDN + 2], w2[FDN + 1], w1[FDN + 1];
    fact[0] = 1;
    for (int k = 1; k < 2 * FDN + 2; ++k) fact[k] = fact[k - 1] * k;
    w2[0] = w1[0] = 0;
    for (int k = 1; k <= FDN; ++k) {  // src/initialization.m:403-419
      const double c = fact[FDN] * fact[FDN] / (fact[FDN - k] * fact[FDN + k]);
      w2[k] = 2 * ((k % 2) ? 1.0 : -1.0) * c / (k * k);
      w1[k] = ((k % 2) ? 1.0 : -1.0) * c / k;
      w2[0] -= 2.0 / (k * k);
    }
    for (int k = 0; k <= FDN; ++k) {
      P.w2[k] = w2[k];
      P.wz2[k] = w2[k] / (ps->g.dz * ps->g.dz);
      P.wt1[k] = -2.0 * ps->g.twist * w1[k] / ps->g.dt;
      P.wz1[k] = w1[k] / ps->g.dz;
    }
  }
  return P;
}

// y = Lap_std x
__global__ void __launch_bounds__(PB) apply_kernel(PoissonOp P, const double* __restrict__ x,
                                                   const double* __restrict__ T, double* __restrict__ y) {
  const Pt q = point_of(P);
  if (!q.ok) return;
  y[q.n] = lap_at(P, x, T, q);
}

int poisson_apply(cyx_poisson_s* ps, const double* x, double* y) {
  const int64_t N = ps->N();
  hipStream_t st = ps->ctx->stream;
  const PoissonOp P = make_op(ps);
  const int nblk = (int)ceil_div(N, PB);
  CYX_CHECK(ensure(ps->tbuf, (size_t)N * sizeof(double)) == CYX_OK, "allocation failed");
  double* T = (double*)ps->tbuf.p;
  if (P.helical) {
    const ThetaShape ts = theta_shape(P, x, T);
    hipLaunchKernelGGL(theta_kernel, dim3(ts.ntb), dim3(PB), 0, st, P, x, T, ts.ntb, ts.pair, FinArgs{});
  }
  hipLaunchKernelGGL(apply_kernel, dim3(nblk), dim3(PB), 0, st, P, x, (const double*)T, y);
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

static int run_impl(cyx_poisson_s* ps, const double* f, double* x, int use_guess, double tol, int max_iter,
                    double omega, double beta, int m, int p, cyx_aar_info* info, cyx_unet_s* net, bool* diverged) {
  *diverged = false;
  CYX_CHECK(m >= 1 && m <= 8, "AAR history m must be in [1, 8]");
  CYX_CHECK(p >= 1 && max_iter >= 1 && tol > 0, "invalid AAR parameters");
  const int64_t N = ps->N();
  CYX_CHECK(N < ((int64_t)1 << 31) - PB, "grid too large for 32-bit point indices");
  hipStream_t st = ps->ctx->stream;
  const int nblk = (int)ceil_div(N, PB);
  const int nblk_norm = (int)std::min<int64_t>(ceil_div(N, AB), 1024);
  const int nstep = step_blocks(make_op(ps));
  const int ndmax = 1 + m * (m + 1) / 2 + m;
  // scratch: xalt | xalt2 | res0 | res1 | T | DX | DF | partials x 2 | state | [r1 | u]
  const size_t vec = (size_t)N * sizeof(double);
  const int64_t pstride = (int64_t)std::max({nblk, nblk_norm, nstep}) * ndmax;  // doubles per partial buffer
  const size_t state_end = vec * (5 + 2 * (size_t)m) + 2 * (size_t)pstride * sizeof(double) + sizeof(AarState) + 256;
  const size_t need = state_end + (net ? 2 * vec + (size_t)N * sizeof(float) + 256 : 0);
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
  double* xalt2 = xalt + N;
  double* R[2] = {xalt + 2 * N, xalt + 3 * N};
  double* T = xalt + 4 * N;
  double* DX = xalt + 5 * N;
  double* DF = DX + (int64_t)m * N;
  double* part = DF + (int64_t)m * N;
  double* parts[2] = {part, part + pstride};  // partials of odd / even iterations
  AarState* S = (AarState*)(((uintptr_t)(part + 2 * pstride) + 255) & ~(uintptr_t)255);
  double* r1buf = net ? (double*)(((uintptr_t)ps->scratch + state_end + 255) & ~(uintptr_t)255) : nullptr;
  double* dbuf = net ? r1buf + N : nullptr;
  float* ubuf = net ? (float*)(dbuf + N) : nullptr;
  // iterate c lives in X[c % 3] (X[1] = the caller's x, holding the guess):
  // a step writes X[(c + 1) % 3] while the finalise of step c - 1, which may
  // select X[(c - 1) % 3], runs in the same launch
  double* X[3] = {xalt, x, xalt2};
  if (!use_guess) hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, x, N, 1.0);
  CYX_HIP(hipMemsetAsync(DX, 0, vec * 2 * m, st));
  CYX_HIP(hipMemsetAsync(T, 0, vec, st));
  CYX_HIP(hipMemsetAsync(S, 0, sizeof(AarState), st));
  hipLaunchKernelGGL(norm_kernel, dim3(nblk_norm), dim3(AB), 0, st, f, N, part);
  hipLaunchKernelGGL(norm_finalize_kernel, dim3(1), dim3(AB), 0, st, (const double*)part, nblk_norm, S);
  const PoissonOp P = make_op(ps);
  AarState hs{};
  PhaseTimer tm(ps->ctx, PH_AAR);
  int count = 1;
  // a Richardson iteration whose finalise has not been enqueued yet: it
  // rides on the next step launch, or runs alone before a poll
  bool fin_pending = false;
  auto fin_args = [&](int c) { return FinArgs{(const double*)parts[c & 1], nstep, c, max_iter, c % 3, tol, S}; };
  // iterations enqueued between polls of the device flag (the network's
  // kernels do not test it, so the learned path polls every 6).  A poll costs
  // an idle gap of tens of microseconds, an iteration enqueued past
  // convergence a no-op launch of ~16 us, so the chunk follows the predicted
  // remaining count: log-residual slope between the last two polls, 8..64
  // iterations, 32 until a slope is known.
  int chunk = net ? 6 : 32;
  double prev_lr = 0.0;
  int prev_it = 0;
  bool have_prev = false;
  while (count <= max_iter) {
    const int cend = std::min(max_iter, count + chunk - 1);
    for (; count <= cend; ++count) {
      const int xb = count & 1;
      StepArgs A{};
      A.f = f;
      A.xp = X[count % 3];
      A.xold = X[(count + 2) % 3];
      A.xnew = X[(count + 1) % 3];
      A.res = R[xb];
      A.fold = R[xb ^ 1];
      A.DX = DX;
      A.DF = DF;
      A.part = parts[xb];
      A.nstep = nstep;
      A.S = S;
      A.hcol = count > 1 ? (count - 2) % m : -1;
      A.anderson = (count % p == 0) ? 1 : 0;
      A.count = count;
      A.max_iter = max_iter;
      A.xbuf = count % 3;
      A.tol = tol;
      A.omega = omega;
      const FinArgs F = fin_args(count - 1);
      const bool fin = fin_pending;
      fin_pending = false;
      if (!A.anderson) {
        hipLaunchKernelGGL(aar_step_brick_kernel<0>, dim3(nstep + (fin ? 1 : 0)), dim3(PB), 0, st, P, A, F,
                           fin ? 1 : 0);
        fin_pending = true;
      } else {
        double* xn = X[(count + 1) % 3];
        switch (m) {
          case 1: launch_anderson<1>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          case 2: launch_anderson<2>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          case 3: launch_anderson<3>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          case 4: launch_anderson<4>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          case 5: launch_anderson<5>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          case 6: launch_anderson<6>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          case 7: launch_anderson<7>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
          default: launch_anderson<8>(nblk, nstep, st, P, A, beta, xn, F, fin); break;
        }
        if (net) {
          // learned correction of the Anderson iterate, with exact line search
          const AarState* cS = S;
          if (P.helical) {
            const ThetaShape tn = theta_shape(P, xn, T);
            hipLaunchKernelGGL(theta_kernel, dim3(tn.ntb), dim3(PB), 0, st, P, (const double*)xn, T, tn.ntb, tn.pair,
                               FinArgs{});
          }
          hipLaunchKernelGGL(resid_kernel, dim3(nblk), dim3(PB), 0, st, P, f, (const double*)xn, (const double*)T,
                             r1buf, cS);
          int rc = unet_precondition(net, ps->coef, r1buf, ubuf);
          if (rc != CYX_OK) return rc;
          hipLaunchKernelGGL(ls_dir_kernel, dim3(nblk), dim3(PB), 0, st, P, (const double*)r1buf, (const float*)ubuf,
                             dbuf, cS);
          if (P.helical) {
            const ThetaShape tn = theta_shape(P, dbuf, T);
            hipLaunchKernelGGL(theta_kernel, dim3(tn.ntb), dim3(PB), 0, st, P, (const double*)dbuf, T, tn.ntb,
                               tn.pair, FinArgs{});
          }
          hipLaunchKernelGGL(ls_apply_kernel, dim3(nblk), dim3(PB), 0, st, P, (const double*)dbuf, (const double*)T,
                             (const double*)r1buf, part, cS);
          hipLaunchKernelGGL(ls_finalize_kernel, dim3(1), dim3(AB), 0, st, (const double*)part, nblk, S);
          hipLaunchKernelGGL(ls_axpy_kernel, dim3(nblk), dim3(PB), 0, st, P, xn, (const double*)dbuf, cS);
        }
      }
    }
    if (fin_pending) {
      hipLaunchKernelGGL(theta_kernel, dim3(1), dim3(PB), 0, st, P, (const double*)X[(count - 1) % 3], T, 0, 0,
                         fin_args(count - 1));
      fin_pending = false;
    }
    CYX_HIP(hipGetLastError());
    CYX_HIP(hipMemcpyAsync(&hs, S, sizeof(AarState), hipMemcpyDeviceToHost, st));
    CYX_HIP(hipStreamSynchronize(st));
    if (hs.done) break;
    // a learned preconditioner that makes the iteration blow up is abandoned
    if (net && !(hs.relres <= 1e3)) {
      *diverged = true;
      return CYX_ERR_NOCONV;
    }
    if (!net && hs.relres > 0.0 && hs.iters > 0) {
      const double lr = std::log(hs.relres);
      chunk = 32;
      if (have_prev && hs.iters > prev_it) {
        const double slope = (lr - prev_lr) / (hs.iters - prev_it);
        if (slope < 0.0) chunk = (int)std::min(64.0, std::max(8.0, std::ceil((std::log(tol) - lr) / slope) + 2.0));
      }
      prev_lr = lr;
      prev_it = hs.iters;
      have_prev = true;
    }
  }
  if (hs.final_buf != 1) CYX_HIP(hipMemcpyAsync(x, X[hs.final_buf], vec, hipMemcpyDeviceToDevice, st));
  CYX_HIP(hipStreamSynchronize(st));
  if (info) {
    info->iterations = hs.iters;
    info->relres = hs.relres;
    info->converged = hs.done == 1;
    info->precond = net ? 1 : 0;
    info->pinv_eig = hs.pinv_eig;
  }
  return hs.done == 1 ? CYX_OK : CYX_ERR_NOCONV;
}

// ---- AAR with the exact inverse of Lap_std as its preconditioner (the place
// of the reference's ILU factors, src/poissonSolve.m:36).  The iteration is
// src/aar.m:33-72 step by step: the residual of x_prev and its norm, the
// preconditioned residual, the history columns, a Richardson or an Anderson
// update; the iterate returned is the one whose residual met the tolerance.
// With M = Lap_std^-1 the preconditioned residual is the error itself, so the
// first Anderson step (count = p) lands on the solution: a solve takes p + 1
// iterations at any tolerance.  The host reads one norm per iteration and,
// at Anderson steps, the m(m+1)/2 + m history products (gamma = pinv(DF'DF)
// DF'res with MATLAB's tolerance, on the host: m <= 8).

// r = f - Lap x; per-block partial sums of r^2
__global__ void __launch_bounds__(PB) sp_resid_kernel(PoissonOp P, const double* __restrict__ f,
                                                      const double* __restrict__ x, const double* __restrict__ T,
                                                      double* __restrict__ r, double* __restrict__ part) {
  __shared__ double red[PB / 64];
  const Pt q = point_of(P);
  double v = 0.0;
  if (q.ok) {
    v = f[q.n] - lap_at(P, x, T, q);
    r[q.n] = v;
  }
  double vv[1] = {v * v};
  const double s = block_sums<1, PB>(vv, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// history columns DX_k = x_prev - Xold, DF_k = res - Fold (hist), then
// Xold = x_prev, Fold = res; Richardson x_new = x_prev + omega res (rich)
__global__ void __launch_bounds__(256) sp_hist_kernel(int64_t N, const double* __restrict__ res,
                                                      const double* __restrict__ xp, double* __restrict__ Xold,
                                                      double* __restrict__ Fold, double* __restrict__ DXk,
                                                      double* __restrict__ DFk, int hist, int rich, double omega,
                                                      double* __restrict__ xn) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const double x = xp[n], rs = res[n];
  if (hist) {
    DXk[n] = x - Xold[n];
    DFk[n] = rs - Fold[n];
  }
  Xold[n] = x;
  Fold[n] = rs;
  if (rich) xn[n] = __builtin_fma(omega, rs, x);
}

// per-block partials of DF_a . DF_b (a <= b, row-major upper triangle) and
// DF_a . res: part[blk * nd + d]
template <int M>
__global__ void __launch_bounds__(256) sp_dots_kernel(int64_t N, const double* __restrict__ DF,
                                                      const double* __restrict__ res, double* __restrict__ part) {
  constexpr int ND = M * (M + 1) / 2 + M;
  __shared__ double red[(256 / 64) * ND];
  double v[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) v[d] = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    double c[M];
#pragma unroll
    for (int a = 0; a < M; ++a) c[a] = DF[(int64_t)a * N + n];
    const double rs = res[n];
    int d = 0;
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
      for (int b = a; b < M; ++b) {
        v[d] = __builtin_fma(c[a], c[b], v[d]);
        ++d;
      }
#pragma unroll
    for (int a = 0; a < M; ++a) {
      v[d] = __builtin_fma(c[a], rs, v[d]);
      ++d;
    }
  }
  const double s = block_sums<ND, 256>(v, red);
  if (threadIdx.x < ND) part[(int64_t)blockIdx.x * ND + threadIdx.x] = s;
}

// x_new = x_prev + beta res - sum_k gamma_k (DX_k + beta DF_k)
__global__ void __launch_bounds__(256) sp_anderson_kernel(int64_t N, int m, const double* __restrict__ xp,
                                                          const double* __restrict__ res,
                                                          const double* __restrict__ DX,
                                                          const double* __restrict__ DF, const double* __restrict__ g,
                                                          double beta, double* __restrict__ xn) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  double acc = __builtin_fma(beta, res[n], xp[n]);
  for (int k = 0; k < m; ++k) acc -= g[k] * __builtin_fma(beta, DF[(int64_t)k * N + n], DX[(int64_t)k * N + n]);
  xn[n] = acc;
}

// gamma = pinv(A) b for the symmetric m x m A (row-major), MATLAB pinv's
// tolerance max(m, m) * norm(A) * eps, by cyclic Jacobi; returns whether an
// eigenvalue was cut
static bool host_pinv_solve(int m, std::vector<double> A, const double* b, double* g) {
  std::vector<double> V((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) V[(size_t)i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) (i == j ? tot : off) += A[(size_t)i * m + j] * A[(size_t)i * m + j];
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[(size_t)p * m + q];
        if (apq == 0.0) continue;
        const double th = 0.5 * (A[(size_t)q * m + q] - A[(size_t)p * m + p]) / apq;
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < m; ++k) {  // A <- J' A J
          const double akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
          A[(size_t)k * m + p] = c * akp - sn * akq;
          A[(size_t)k * m + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
          A[(size_t)p * m + k] = c * apk - sn * aqk;
          A[(size_t)q * m + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = V[(size_t)k * m + p], vkq = V[(size_t)k * m + q];
          V[(size_t)k * m + p] = c * vkp - sn * vkq;
          V[(size_t)k * m + q] = sn * vkp + c * vkq;
        }
      }
  }
  double nrm = 0.0;
  for (int i = 0; i < m; ++i) nrm = std::max(nrm, std::fabs(A[(size_t)i * m + i]));
  const double tol = m * nrm * 2.220446049250313e-16;
  bool cut = false;
  for (int i = 0; i < m; ++i) g[i] = 0.0;
  for (int e = 0; e < m; ++e) {
    const double lam = A[(size_t)e * m + e];
    if (std::fabs(lam) <= tol) {
      cut = true;
      continue;
    }
    double vb = 0.0;
    for (int i = 0; i < m; ++i) vb += V[(size_t)i * m + e] * b[i];
    for (int i = 0; i < m; ++i) g[i] += V[(size_t)i * m + e] * vb / lam;
  }
  return cut;
}

static int run_spectral(cyx_poisson_s* ps, const double* f, double* x, int use_guess, double tol, int max_iter,
                        double omega, double beta, int m, int p, cyx_aar_info* info) {
  CYX_CHECK(m >= 1 && m <= 8, "AAR history m must be in [1, 8]");
  CYX_CHECK(p >= 1 && max_iter >= 1 && tol > 0, "invalid AAR parameters");
  const int64_t N = ps->N();
  CYX_CHECK(N < ((int64_t)1 << 31) - PB, "grid too large for 32-bit point indices");
  hipStream_t st = ps->ctx->stream;
  const PoissonOp P = make_op(ps);
  const int nblk = (int)ceil_div(N, PB);
  const int ndot = 1024;  // blocks of the history products
  const int nd = m * (m + 1) / 2 + m;
  // scratch: xa | xb | r | res | Xold | Fold | T | DX | DF | partials | gamma
  const size_t vec = (size_t)N * sizeof(double);
  const size_t npart = (size_t)std::max(nblk, ndot * nd) + 64;
  const size_t need = vec * (7 + 2 * (size_t)m) + npart * sizeof(double) + 8 * sizeof(double) + 256;
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
  double* xa = ps->scratch;
  double* r = xa + N;
  double* res = r + N;
  double* Xold = res + N;
  double* Fold = Xold + N;
  double* T = Fold + N;
  double* DX = T + N;
  double* DF = DX + (int64_t)m * N;
  double* part = DF + (int64_t)m * N;
  double* gdev = part + npart;
  // x_prev alternates between the caller's x and xa
  double* X[2] = {x, xa};
  const dim3 gv(ceil_div(N, 256));
  if (!use_guess) hipLaunchKernelGGL(fill_kernel, gv, dim3(256), 0, st, x, N, 1.0);
  CYX_HIP(hipMemsetAsync(DX, 0, vec * 2 * m, st));
  CYX_HIP(hipMemsetAsync(T, 0, vec, st));
  std::vector<double> hp(std::max((size_t)nblk, (size_t)ndot * nd));
  // ||f||
  hipLaunchKernelGGL(norm_kernel, dim3(nblk < 1024 ? nblk : 1024), dim3(AB), 0, st, f, N, part);
  const int nnorm = nblk < 1024 ? nblk : 1024;
  CYX_HIP(hipMemcpyAsync(hp.data(), part, (size_t)nnorm * sizeof(double), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  double fn2 = 0.0;
  for (int b = 0; b < nnorm; ++b) fn2 += hp[b];
  CYX_CHECK(fn2 > 0.0, "AAR right-hand side is zero");
  const double nb = 1.0 / std::sqrt(fn2);
  PhaseTimer tm(ps->ctx, PH_AAR);
  int count = 1, cur = 0, pinv_cut = 0;
  double relres = 1.0 + tol;
  std::vector<double> A((size_t)m * m), bv(m), gam(m);
  while (count <= max_iter) {
    double* xp = X[cur];
    double* xn = X[cur ^ 1];
    if (P.helical) {
      const ThetaShape ts = theta_shape(P, xp, T);
      hipLaunchKernelGGL(theta_kernel, dim3(ts.ntb), dim3(PB), 0, st, P, (const double*)xp, T, ts.ntb, ts.pair,
                         FinArgs{});
    }
    hipLaunchKernelGGL(sp_resid_kernel, dim3(nblk), dim3(PB), 0, st, P, f, (const double*)xp, (const double*)T, r,
                       part);
    CYX_HIP(hipGetLastError());
    CYX_HIP(hipMemcpyAsync(hp.data(), part, (size_t)nblk * sizeof(double), hipMemcpyDeviceToHost, st));
    CYX_HIP(hipStreamSynchronize(st));
    double r2 = 0.0;
    for (int b = 0; b < nblk; ++b) r2 += hp[b];
    relres = std::sqrt(r2) * nb;
    if (relres <= tol) break;  // x_prev is the iterate src/aar.m returns
    if (count == max_iter) break;
    int rc = spectral_solve_real(ps->spec, r, res, 1, N, st);
    if (rc != CYX_OK) return rc;
    const int anderson = (count % p == 0) ? 1 : 0;
    const int k = count > 1 ? (count - 2) % m : 0;
    hipLaunchKernelGGL(sp_hist_kernel, gv, dim3(256), 0, st, N, (const double*)res, (const double*)xp, Xold, Fold,
                       DX + (int64_t)k * N, DF + (int64_t)k * N, count > 1 ? 1 : 0, anderson ? 0 : 1, omega, xn);
    if (anderson) {
      switch (m) {
#define CYX_SPD(MM) \
  case MM: hipLaunchKernelGGL(sp_dots_kernel<MM>, dim3(ndot), dim3(256), 0, st, N, (const double*)DF, (const double*)res, part); break;
        CYX_SPD(1) CYX_SPD(2) CYX_SPD(3) CYX_SPD(4) CYX_SPD(5) CYX_SPD(6) CYX_SPD(7) CYX_SPD(8)
#undef CYX_SPD
      }
      CYX_HIP(hipMemcpyAsync(hp.data(), part, (size_t)ndot * nd * sizeof(double), hipMemcpyDeviceToHost, st));
      CYX_HIP(hipStreamSynchronize(st));
      std::vector<double> tot(nd, 0.0);
      for (int b = 0; b < ndot; ++b)
        for (int d = 0; d < nd; ++d) tot[d] += hp[(size_t)b * nd + d];
      int d = 0;
      for (int a = 0; a < m; ++a)
        for (int b2 = a; b2 < m; ++b2) {
          A[(size_t)a * m + b2] = tot[d];
          A[(size_t)b2 * m + a] = tot[d];
          ++d;
        }
      for (int a = 0; a < m; ++a) bv[a] = tot[d++];
      pinv_cut += host_pinv_solve(m, A, bv.data(), gam.data()) ? 1 : 0;
      CYX_HIP(hipMemcpyAsync(gdev, gam.data(), (size_t)m * sizeof(double), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(sp_anderson_kernel, gv, dim3(256), 0, st, N, m, (const double*)xp, (const double*)res,
                         (const double*)DX, (const double*)DF, (const double*)gdev, beta, xn);
    }
    CYX_HIP(hipGetLastError());
    cur ^= 1;
    ++count;
  }
  if (X[cur] != x) CYX_HIP(hipMemcpyAsync(x, X[cur], vec, hipMemcpyDeviceToDevice, st));
  CYX_HIP(hipStreamSynchronize(st));
  const bool conv = relres <= tol;
  if (info) {
    info->iterations = count;
    info->relres = relres;
    info->converged = conv ? 1 : 0;
    info->precond = 3;
    info->pinv_eig = pinv_cut;
  }
  return conv ? CYX_OK : CYX_ERR_NOCONV;
}

int poisson_run(cyx_poisson_s* ps, const double* f, double* x, int use_guess, double tol, int max_iter,
                double omega, double beta, int m, int p, cyx_aar_info* info) {
  bool diverged = false;
  cyx_unet_s* net = ps->net;
  if (ps->spec && ps->use_spectral) return run_spectral(ps, f, x, use_guess, tol, max_iter, omega, beta, m, p, info);
  if (!net) return run_impl(ps, f, x, use_guess, tol, max_iter, omega, beta, m, p, info, nullptr, &diverged);
  // keep the guess: a diverging learned solve is restarted with Jacobi
  const int64_t N = ps->N();
  hipStream_t st = ps->ctx->stream;
  CYX_CHECK(ensure(ps->tbuf, 2 * (size_t)N * sizeof(double) + 1024) == CYX_OK, "allocation failed");
  double* gbak = (double*)ps->tbuf.p + N + 128;
  if (use_guess) CYX_HIP(hipMemcpyAsync(gbak, x, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, st));
  int rc = run_impl(ps, f, x, use_guess, tol, max_iter, omega, beta, m, p, info, net, &diverged);
  if (!diverged) return rc;
  if (use_guess) CYX_HIP(hipMemcpyAsync(x, gbak, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, st));
  rc = run_impl(ps, f, x, use_guess, tol, max_iter, omega, beta, m, p, info, nullptr, &diverged);
  if (info) info->precond = 2;
  return rc;
}

}  // namespace cyx
