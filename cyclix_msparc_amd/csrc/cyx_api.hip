// extern "C" entry points of the Cyclix CheFSI / AAR library.  Host-side
// set-up (operator coefficients, KB tables), orchestration of the device
// kernels, and the FP64 dense step of Rayleigh-Ritz via rocSOLVER.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cyx_internal.hpp"

namespace cyx {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int ensure(DeviceBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return CYX_OK;
  if (b.p) {
    (void)hipDeviceSynchronize();
    (void)hipFree(b.p);
  }
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    set_error("device allocation of %zu bytes failed", bytes);
    return CYX_ERR_NOMEM;
  }
  b.bytes = bytes;
  return CYX_OK;
}

size_t dtype_size(int dtype) {
  switch (dtype) {
    case CYX_F64:
      return 8;
    case CYX_C128:
      return 16;
    case CYX_F32:
      return 4;
    case CYX_C64:
      return 8;
  }
  return 0;
}

static bool dtype_complex(int dtype) { return dtype == CYX_C128 || dtype == CYX_C64; }

// FD weights, src/initialization.m:403-419
static void fd_weights(int fdn, double* w2, double* w1) {
  double fact[64];
  fact[0] = 1;
  for (int k = 1; k < 2 * fdn + 2; ++k) fact[k] = fact[k - 1] * k;
  w2[0] = 0;
  w1[0] = 0;
  for (int k = 1; k <= fdn; ++k) {
    const double c = fact[fdn] * fact[fdn] / (fact[fdn - k] * fact[fdn + k]);
    w2[k] = 2 * ((k % 2) ? 1.0 : -1.0) * c / (k * k);
    w1[k] = ((k % 2) ? 1.0 : -1.0) * c / k;
    w2[0] -= 2.0 / (k * k);
  }
}

// stencil = true: the grid a Hamiltonian handle runs the marching stencil on
// (theta rings up to MAX_NT nodes); the Poisson operator reads its theta taps
// from memory and takes any ring
static int check_grid(const cyx_grid_desc* g, bool stencil = true) {
  CYX_CHECK(g != nullptr, "grid is null");
  CYX_CHECK(g->fdn == FDN, "only FDn = %d is supported (got %d)", FDN, g->fdn);
  CYX_CHECK(g->nr >= 1 && g->nt >= FDN && g->nz >= FDN, "grid %dx%dx%d too small for FDn=%d", g->nr, g->nt, g->nz,
            FDN);
  CYX_CHECK(!stencil || g->nt <= MAX_NT, "nt = %d exceeds %d", g->nt, MAX_NT);
  CYX_CHECK(g->dr > 0 && g->dt > 0 && g->dz > 0, "spacings must be positive");
  CYX_CHECK(g->cell_typ == 3 || g->cell_typ == 4 || g->cell_typ == 5, "cell_typ must be 3, 4 or 5");
  CYX_CHECK(g->cell_typ == 3 || g->twist != 0.0, "helical cell_typ %d needs a non-zero twist", g->cell_typ);
  CYX_CHECK(g->xin > 0.0, "inner radius xin must be positive");
  return CYX_OK;
}

// Per-radius operator coefficients with -1/2 folded in (scale = -0.5), or
// the plain Lap_std coefficients (scale = 1).  Layout per radius:
// [c+_1..c+_6, c-_1..c-_6, c_theta, c_diag]
static std::vector<double> radial_coef(const cyx_grid_desc& g, double scale) {
  double w2[FDN + 1], w1[FDN + 1];
  fd_weights(FDN, w2, w1);
  const double tau2 = (g.cell_typ == 3) ? 0.0 : g.twist * g.twist;
  std::vector<double> c((size_t)g.nr * RCOEF);
  for (int i = 0; i < g.nr; ++i) {
    const double r = g.xin + i * g.dr;
    double* ci = &c[(size_t)i * RCOEF];
    for (int p = 1; p <= FDN; ++p) {
      ci[p - 1] = scale * (w2[p] / (g.dr * g.dr) + w1[p] / (r * g.dr));
      ci[FDN + p - 1] = scale * (w2[p] / (g.dr * g.dr) - w1[p] / (r * g.dr));
    }
    const double cth = (tau2 + 1.0 / (r * r)) / (g.dt * g.dt);
    ci[2 * FDN] = scale * cth;
    ci[2 * FDN + 1] = scale * (w2[0] / (g.dr * g.dr) + w2[0] * cth + w2[0] / (g.dz * g.dz));
  }
  return c;
}

template <class R>
static int upload_coef(OpCoef<R>& C, const cyx_grid_desc& g, const std::vector<double>& rc, const double* veff,
                       int64_t N) {
  double w2[FDN + 1], w1[FDN + 1];
  fd_weights(FDN, w2, w1);
  const double tau = (g.cell_typ == 3) ? 0.0 : g.twist;
  for (int p = 0; p <= FDN; ++p) {
    C.w2[p] = (R)w2[p];
    C.w1[p] = (R)w1[p];
  }
  C.s_r2 = (R)(-0.5 / (g.dr * g.dr));
  C.s_r1 = (R)(-0.5 / g.dr);
  C.s_z2 = (R)(-0.5 / (g.dz * g.dz));
  C.s_x = (R)((-2.0 * tau / g.dt) * (-0.5 / g.dz));
  std::vector<R> rcr(rc.begin(), rc.end());
  CYX_HIP(hipMalloc((void**)&C.rcoef, rcr.size() * sizeof(R)));
  CYX_HIP(hipMemcpy(C.rcoef, rcr.data(), rcr.size() * sizeof(R), hipMemcpyHostToDevice));
  CYX_HIP(hipMalloc((void**)&C.veff, (size_t)N * sizeof(R)));
  std::vector<R> v(veff, veff + N);
  CYX_HIP(hipMemcpy(C.veff, v.data(), (size_t)N * sizeof(R), hipMemcpyHostToDevice));
  return CYX_OK;
}

template <class T>
struct HostKb {
  using R = typename traits<T>::real;
  std::vector<int> g_ptr, g_node, s_ptr, s_il, il_proj, pr_lo, pr_hi, blk_pair, pp_ptr;
  std::vector<uint32_t> s_pk;
  int max_count = 0;
  std::vector<T> g_coef, il_phase;
  std::vector<R> gamma, s_chi;
};

template <class T>
static T mk(double re, double im) {
  return phase_of<T>(re, im);
}

// Flatten S.Atom(J).rcImage(img) into gather (by projector) and scatter (CSR
// by node) tables, src/h_nonlocal_vector_mult.m:30-43 with the Bloch phase
// exp(i k.(x_J - x_img)) on the gather and its conjugate on the scatter.
template <class T>
static void build_kb(const cyx_projector_desc* d, const double* W, const double kpt[3], int64_t N, HostKb<T>& K) {
  using R = typename traits<T>::real;
  std::vector<int> pbase(d->n_atom + 1, 0);
  for (int a = 0; a < d->n_atom; ++a) pbase[a + 1] = pbase[a] + d->angnum[a];
  const int nproj = pbase[d->n_atom];
  K.gamma.resize(nproj);
  for (int p = 0; p < nproj; ++p) K.gamma[p] = (R)d->gamma[p];
  struct Ent {
    int node, il;
    double chi, gr, gi;
  };
  std::vector<std::vector<Ent>> byproj(nproj);
  std::vector<int> cnt(N + 1, 0);
  int64_t off = 0, choff = 0;
  for (int im = 0; im < d->n_image; ++im) {
    const int a = d->image_atom[im];
    const int nn = d->image_nnodes[im];
    const double* dd = d->image_disp + 3 * im;
    const double ph = kpt[0] * dd[0] + kpt[1] * dd[1] + kpt[2] * dd[2];
    const double c = std::cos(ph), s = std::sin(ph);
    for (int l = 0; l < d->angnum[a]; ++l) {
      const int p = pbase[a] + l;
      const int il = (int)K.il_proj.size();
      K.il_proj.push_back(p);
      K.il_phase.push_back(mk<T>(c, -s));
      for (int e = 0; e < nn; ++e) {
        const int node = d->rc_pos[off + e];
        const double chi = d->chi[choff + (int64_t)l * nn + e];
        const double gw = chi * W[node];
        byproj[p].push_back(Ent{node, il, chi, gw * c, gw * s});
        cnt[node + 1]++;
      }
    }
    off += nn;
    choff += (int64_t)nn * d->angnum[a];
  }
  // each projector's entries ascend by node (images interleave), so that a
  // node range is one contiguous entry range per projector
  for (int p = 0; p < nproj; ++p)
    std::stable_sort(byproj[p].begin(), byproj[p].end(),
                     [](const Ent& x, const Ent& y) { return x.node != y.node ? x.node < y.node : x.il < y.il; });
  K.g_ptr.assign(nproj + 1, 0);
  for (int p = 0; p < nproj; ++p) {
    K.g_ptr[p + 1] = K.g_ptr[p] + (int)byproj[p].size();
    for (const Ent& e : byproj[p]) {
      K.g_node.push_back(e.node);
      K.g_coef.push_back(mk<T>(e.gr, e.gi));
    }
  }
  K.s_ptr.assign(N + 1, 0);
  for (int64_t n = 0; n < N; ++n) K.s_ptr[n + 1] = K.s_ptr[n] + cnt[n + 1];
  K.s_chi.assign(K.s_ptr[N], R(0));
  K.s_il.assign(K.s_ptr[N], 0);
  std::vector<int> fill(K.s_ptr.begin(), K.s_ptr.end() - 1);
  for (int p = 0; p < nproj; ++p)
    for (const Ent& e : byproj[p]) {
      const int k = fill[e.node]++;
      K.s_chi[k] = (R)e.chi;
      K.s_il[k] = e.il;
    }
  // one word per node for the stencil: (first entry << 6) | count
  K.s_pk.assign(N, 0u);
  for (int64_t n = 0; n < N; ++n) {
    const int c = K.s_ptr[n + 1] - K.s_ptr[n];
    K.max_count = std::max(K.max_count, c);
    K.s_pk[n] = ((uint32_t)K.s_ptr[n] << 6) | (uint32_t)std::min(c, 63);
  }
  // gather chunks: node ranges holding about KB_CHUNK_ENTRIES entries of all
  // projectors together (cnt[n + 1] is node n's entry count)
  constexpr int KB_CHUNK_ENTRIES = 2048;
  std::vector<int64_t> bound{0};
  for (int64_t n = 0, acc = 0; n < N; ++n) {
    acc += cnt[n + 1];
    if (acc >= KB_CHUNK_ENTRIES && n + 1 < N) {
      bound.push_back(n + 1);
      acc = 0;
    }
  }
  bound.push_back(N);
  const int nch = (int)bound.size() - 1;
  // pair ids run projector-major, in chunk order within a projector
  std::vector<std::vector<int>> by_chunk(nch);  // pair ids of each chunk, by projector
  K.pp_ptr.assign(nproj + 1, 0);
  for (int p = 0; p < nproj; ++p) {
    const int e0 = K.g_ptr[p], e1 = K.g_ptr[p + 1];
    int e = e0;
    for (int c = 0; c < nch && e < e1; ++c) {
      const int lo = e;
      while (e < e1 && K.g_node[e] < bound[c + 1]) ++e;
      if (e > lo) {
        const int q = (int)K.pr_lo.size();
        K.pr_lo.push_back(lo);
        K.pr_hi.push_back(e);
        by_chunk[c].push_back(q);
      }
    }
    K.pp_ptr[p + 1] = (int)K.pr_lo.size();
  }
  // pairs of chunk c go to launch slots congruent to c mod 8 (one XCD)
  std::vector<std::vector<int>> slot(8);
  for (int c = 0; c < nch; ++c)
    for (int q : by_chunk[c]) slot[c % 8].push_back(q);
  size_t len = 0;
  for (const auto& v : slot) len = std::max(len, v.size());
  K.blk_pair.assign(8 * len, -1);
  for (int x = 0; x < 8; ++x)
    for (size_t i = 0; i < slot[x].size(); ++i) K.blk_pair[8 * i + x] = slot[x][i];
}

template <class V>
static int upload_vec(V*& dst, const std::vector<V>& src) {
  if (src.empty()) {
    dst = nullptr;
    return CYX_OK;
  }
  CYX_HIP(hipMalloc((void**)&dst, src.size() * sizeof(V)));
  CYX_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
  return CYX_OK;
}

template <class T>
static int upload_kb(KbTables<T>& D, const cyx_projector_desc* d, const double* W, const double kpt[3], int64_t N) {
  HostKb<T> K;
  build_kb<T>(d, W, kpt, N, K);
  CYX_CHECK(K.max_count <= 63, "a grid node lies in %d projector balls (at most 63 supported)", K.max_count);
  CYX_CHECK(K.s_chi.size() < (size_t)1 << 26, "too many projector entries (%zu)", K.s_chi.size());
  D.nproj = (int)K.gamma.size();
  D.nimgl = (int)K.il_proj.size();
  int rc = CYX_OK;
  if ((rc = upload_vec(D.g_ptr, K.g_ptr))) return rc;
  if ((rc = upload_vec(D.g_node, K.g_node))) return rc;
  if ((rc = upload_vec(D.g_coef, K.g_coef))) return rc;
  if ((rc = upload_vec(D.gamma, K.gamma))) return rc;
  if ((rc = upload_vec(D.s_pk, K.s_pk))) return rc;
  if ((rc = upload_vec(D.s_chi, K.s_chi))) return rc;
  if ((rc = upload_vec(D.s_il, K.s_il))) return rc;
  if ((rc = upload_vec(D.il_phase, K.il_phase))) return rc;
  if ((rc = upload_vec(D.il_proj, K.il_proj))) return rc;
  D.npair = (int)K.pr_lo.size();
  D.nblk_pair = (int)K.blk_pair.size();
  if ((rc = upload_vec(D.pr_lo, K.pr_lo))) return rc;
  if ((rc = upload_vec(D.pr_hi, K.pr_hi))) return rc;
  if ((rc = upload_vec(D.blk_pair, K.blk_pair))) return rc;
  if ((rc = upload_vec(D.pp_ptr, K.pp_ptr))) return rc;
  {
    std::vector<int> nodes;
    for (int64_t n = 0; n < (int64_t)K.s_pk.size(); ++n)
      if (K.s_pk[n] & 63u) nodes.push_back((int)n);
    D.nkbn = (int)nodes.size();
    if ((rc = upload_vec(D.kb_nodes, nodes))) return rc;
    std::vector<T> cph(K.s_chi.size());
    std::vector<int> pj(K.s_chi.size());
    for (size_t e = 0; e < cph.size(); ++e) {
      const T ph = K.il_phase[K.s_il[e]];
      if constexpr (traits<T>::is_complex) {
        cph[e].x = K.s_chi[e] * ph.x;
        cph[e].y = K.s_chi[e] * ph.y;
      } else {
        cph[e] = K.s_chi[e] * ph;
      }
      pj[e] = K.il_proj[K.s_il[e]];
    }
    if ((rc = upload_vec(D.s_cph, cph))) return rc;
    if ((rc = upload_vec(D.s_pj, pj))) return rc;
  }
  return CYX_OK;
}

template <class T>
static void free_kb(KbTables<T>& D) {
  (void)hipFree(D.g_ptr);
  (void)hipFree(D.g_node);
  (void)hipFree(D.g_coef);
  (void)hipFree(D.gamma);
  (void)hipFree(D.s_pk);
  (void)hipFree(D.s_chi);
  (void)hipFree(D.s_il);
  (void)hipFree(D.il_phase);
  (void)hipFree(D.il_proj);
  (void)hipFree(D.kb_nodes);
  (void)hipFree(D.s_cph);
  (void)hipFree(D.s_pj);
  (void)hipFree(D.pr_lo);
  (void)hipFree(D.pr_hi);
  (void)hipFree(D.blk_pair);
  (void)hipFree(D.pp_ptr);
  D = KbTables<T>{};
}

static int check_dtype(cyx_ham_s* h, int dtype) {
  CYX_CHECK(dtype >= CYX_F64 && dtype <= CYX_C64, "unknown dtype %d", dtype);
  CYX_CHECK(!h->bloch || dtype_complex(dtype), "a Bloch channel (non-zero kpt) needs a complex dtype");
  return CYX_OK;
}

// ---- small device helpers for Lanczos (vectors of T, FP64 dot products)
template <class T>
__global__ void lz_dot_kernel(const T* a, const T* b, const double* W, int64_t N, double* part) {
  __shared__ double sr[256], si[256];
  double re = 0, im = 0;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const double w = W[n];
    if constexpr (traits<T>::is_complex) {
      re += w * (a[n].x * b[n].x + a[n].y * b[n].y);
      im += w * (a[n].x * b[n].y - a[n].y * b[n].x);
    } else {
      re += w * a[n] * b[n];
    }
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr[0];
    part[2 * blockIdx.x + 1] = si[0];
  }
}

// w = w - alpha v - beta vprev, then (optionally) vprev = v, v = w / nrm
template <class T>
__global__ void lz_axpy_kernel(T* w, const T* v, const T* vprev, int64_t N, double alpha, double beta) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  w[n] = rfma(-beta, vprev[n], rfma(-alpha, v[n], w[n]));
}
template <class T>
__global__ void lz_scale_kernel(T* dst, const T* src, int64_t N, double s) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  dst[n] = rmul(s, src[n]);
}
template <class T>
__global__ void lz_load_kernel(T* dst, const double* v0, int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  if constexpr (traits<T>::is_complex)
    dst[n] = T{v0[n], 0.0};
  else
    dst[n] = v0[n];
}

// extreme eigenvalues of the symmetric tridiagonal (alpha, beta) by bisection
static double tri_extreme(const std::vector<double>& a, const std::vector<double>& b, bool want_max) {
  const int n = (int)a.size();
  double lo = a[0], hi = a[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0) + (i < n - 1 ? std::fabs(b[i]) : 0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count_below = [&](double x) {
    int c = 0;
    double d = 1.0;
    for (int i = 0; i < n; ++i) {
      d = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++c;
    }
    return c;
  };
  const int target = want_max ? n - 1 : 0;  // index of the wanted eigenvalue
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) > target)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

template <class T>
static int lanczos_impl(cyx_ham_s* h, int dtype, const double* v0, int maxit, double tol, double* lmin, double* lmax) {
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  const int nb = 256;
  size_t need = 3 * (size_t)N * sizeof(T) + 2 * nb * sizeof(double) + (size_t)N * sizeof(double);
  DeviceBuf buf;
  if (ensure(buf, need) != CYX_OK) return CYX_ERR_NOMEM;
  T* v = (T*)buf.p;
  T* vp = v + N;
  T* w = vp + N;
  double* part = (double*)(w + N);
  double* dv0 = part + 2 * nb;
  auto cleanup = [&]() { (void)hipFree(buf.p); };
  const int g256 = ceil_div(N, 256);
  if (hipMemcpyAsync(dv0, v0, (size_t)N * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) {
    cleanup();
    set_error("v0 upload failed");
    return CYX_ERR_HIP;
  }
  hipLaunchKernelGGL(lz_load_kernel<T>, dim3(g256), dim3(256), 0, st, v, dv0, N);
  (void)hipMemsetAsync(vp, 0, (size_t)N * sizeof(T), st);
  std::vector<double> hp(2 * nb);
  auto dot = [&](const T* a, const T* b) -> double {
    hipLaunchKernelGGL(lz_dot_kernel<T>, dim3(nb), dim3(256), 0, st, a, b, (const double*)h->Wm, N, part);
    (void)hipMemcpyAsync(hp.data(), part, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    double s = 0;
    for (int k = 0; k < nb; ++k) s += hp[2 * k];
    return s;
  };
  double nrm = std::sqrt(dot(v, v));
  hipLaunchKernelGGL(lz_scale_kernel<T>, dim3(g256), dim3(256), 0, st, v, (const T*)v, N, 1.0 / nrm);
  std::vector<double> al, be;
  double beta = 0, lo = 0, hi = 0, lo_prev = 0, hi_prev = 0;
  int rc = CYX_OK;
  for (int j = 0; j < maxit; ++j) {
    if ((rc = stencil_step(h, dtype, v, nullptr, w, 1, N, 1.0, 0.0, 0.0)) != CYX_OK) break;
    const double alpha = dot(v, w);
    hipLaunchKernelGGL(lz_axpy_kernel<T>, dim3(g256), dim3(256), 0, st, w, (const T*)v, (const T*)vp, N, alpha, beta);
    al.push_back(alpha);
    beta = std::sqrt(dot(w, w));
    be.push_back(beta);
    std::vector<double> bb(be.begin(), be.end() - 1);
    lo = tri_extreme(al, bb, false);
    hi = tri_extreme(al, bb, true);
    if (j >= 4 && std::fabs(lo - lo_prev) <= tol * std::max(1.0, std::fabs(lo)) &&
        std::fabs(hi - hi_prev) <= tol * std::max(1.0, std::fabs(hi)))
      break;
    lo_prev = lo;
    hi_prev = hi;
    if (beta == 0.0) break;
    // vp = v; v = w / beta
    (void)hipMemcpyAsync(vp, v, (size_t)N * sizeof(T), hipMemcpyDeviceToDevice, st);
    hipLaunchKernelGGL(lz_scale_kernel<T>, dim3(g256), dim3(256), 0, st, v, (const T*)w, N, 1.0 / beta);
  }
  (void)hipStreamSynchronize(st);
  cleanup();
  if (rc != CYX_OK) return rc;
  *lmin = lo;
  *lmax = hi;
  return CYX_OK;
}

// ---- Rayleigh-Ritz on psi (filtered), O = other buffer; result ends in O
static int rr_impl(cyx_ham_s* h, int dtype, void* psi, void* other, int ns, double* eigval, void** result) {
  const int64_t N = h->N();
  hipStream_t st = h->ctx->stream;
  const bool cx = dtype_complex(dtype);
  const size_t esz = cx ? 16 : 8;
  const size_t mat = (size_t)ns * ns * esz;
  // tridiagonal reduction: distributed over workgroups with the matrix in LDS,
  // else one workgroup, else rocSOLVER's heevd (CYX_TRIDIAG = dist | one |
  // lapack forces a path where it fits; read per call)
  const char* trd_env = getenv("CYX_TRIDIAG");
  const std::string trd_req = trd_env ? trd_env : "";
  const bool dist_fits = hetrd_dist_groups(ns, cx) > 0, one_fits = hetrd_lds_bytes(ns, cx) <= 64 * 1024;
  const bool use_dist = dist_fits && (trd_req.empty() || trd_req == "dist");
  const bool own_trd = use_dist || (one_fits && trd_req != "lapack");
  const size_t dist_ws = use_dist ? hetrd_dist_ws_bytes(ns, cx) : 0;
  // FP32 blocks: Gram and projection on the FP32 matrix cores (FP64 chunk
  // sums), rotation there too when the overlap is well conditioned
  // (CYX_SUBSPACE = fp64 keeps the FP64 matrix-core products; read per call)
  const char* sub_env = getenv("CYX_SUBSPACE");
  const bool fp32_block = dtype == CYX_F32 || dtype == CYX_C64;
  const bool sub32 = fp32_block && !(sub_env && std::string(sub_env) == "fp64");
  const size_t need =
      3 * mat + 4 * (size_t)ns * sizeof(double) + (size_t)ns * esz + 512 + dist_ws + 256 + ((size_t)ns + 32) * 8;
  int rc;
  if ((rc = ensure(h->sub, need)) != CYX_OK) return rc;
  char* base = (char*)h->sub.p;
  void* M = base;
  void* Hs = base + mat;
  void* X = base + 2 * mat;  // eigenvectors when the reduction runs in one workgroup
  double* D = (double*)(base + 3 * mat);
  double* E = D + ns;
  double* scal = E + ns;
  void* tau = (void*)(base + 3 * mat + (((size_t)3 * ns * sizeof(double) + 15) & ~(size_t)15));
  rocblas_int* info = (rocblas_int*)((char*)tau + (size_t)ns * esz + 64);
  void* trd_ws = (void*)(((uintptr_t)(info + 16) + 255) & ~(uintptr_t)255);
  unsigned* trd_flag = nullptr;
  // saved diagonal of M and the Cholesky conditioning ratio (after the tridiagonal workspace)
  double* mdiag = (double*)(((uintptr_t)trd_ws + dist_ws + 255) & ~(uintptr_t)255);
  double* chol_ratio_dev = mdiag + ns;
  auto reduce = [&](void* a) -> int {
    return use_dist ? hetrd_dist(st, cx, a, ns, D, E, tau, trd_ws, &trd_flag) : hetrd(st, cx, a, ns, D, E, tau);
  };
  // FP32 blocks: the overlap first, on the FP32 matrix cores; its Cholesky
  // factor either certifies the block well conditioned (min_i R_ii^2 / M_ii
  // >= 1e-2: every column keeps a tenth of its norm outside the span of the
  // earlier ones), and the projection and the rotation stay on the FP32
  // cores, or not, and all three products run on the FP64 cores, whose
  // overlap of the stored block is exact to FP64 rounding
  bool use32 = false, factored = false;
  if (sub32) {
    rocblas_handle bh = (rocblas_handle)h->ctx->blas;
    if ((rc = gram_rect32(h, dtype, psi, N, ns, psi, N, ns, h->Wm, M, true)) != CYX_OK) return rc;
    if ((rc = hermitize(h, dtype, M, ns)) != CYX_OK) return rc;
    if ((rc = overlap_diag(h, cx, M, ns, mdiag)) != CYX_OK) return rc;
    rocblas_status s0;
    {
      PhaseTimer tm(h->ctx, PH_DENSE_EIG);
      s0 = cx ? rocsolver_zpotrf(bh, rocblas_fill_upper, ns, (rocblas_double_complex*)M, ns, info)
              : rocsolver_dpotrf(bh, rocblas_fill_upper, ns, (double*)M, ns, info);
    }
    if (s0 != rocblas_status_success) {
      set_error("overlap Cholesky failed with rocSOLVER status %d", (int)s0);
      return CYX_ERR_SOLVER;
    }
    if ((rc = cholesky_ratio(h, cx, M, mdiag, ns, chol_ratio_dev)) != CYX_OK) return rc;
    double ratio = 0.0;
    rocblas_int pinfo = 0;
    CYX_HIP(hipMemcpyAsync(&ratio, chol_ratio_dev, sizeof(double), hipMemcpyDeviceToHost, st));
    CYX_HIP(hipMemcpyAsync(&pinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
    CYX_HIP(hipStreamSynchronize(st));
    use32 = pinfo == 0 && ratio >= 1e-2;
    factored = use32;
    h->rot32_count += use32 ? 1 : 0;
    h->rot64_count += use32 ? 0 : 1;
  }
  if (!use32)
    if ((rc = gram_w(h, dtype, psi, psi, ns, N, M)) != CYX_OK) return rc;
  // H psi in FP64 arithmetic for every storage type: the FP32 operator's
  // rounded coefficients shift the Ritz values by about eps32 ||H||
  if ((rc = stencil_apply_fp64(h, dtype, psi, other, ns, N)) != CYX_OK) return rc;
  // Hs in full: its two triangles are averaged below, as src/eigSolver.m:108 does
  if (use32) {
    if ((rc = gram_rect32(h, dtype, psi, N, ns, other, N, ns, h->Wm, Hs, false)) != CYX_OK) return rc;
  } else {
    if ((rc = gram_rect(h, dtype, psi, N, ns, other, N, ns, h->Wm, Hs, false)) != CYX_OK) return rc;
  }
  if (!factored)
    if ((rc = hermitize(h, dtype, M, ns)) != CYX_OK) return rc;
  if ((rc = hermitize(h, dtype, Hs, ns)) != CYX_OK) return rc;
  void* Q = Hs;  // eigenvectors of the pencil
  {
    // Hermitian-definite pencil (Hs, M) in FP64: M = U^H U, Hs' = U^-H Hs U^-1,
    // Hs' = Q' D Q'^H, Q = U^-1 Q' (Q^H M Q = I).
    PhaseTimer tm(h->ctx, PH_DENSE_EIG);
    rocblas_handle bh = (rocblas_handle)h->ctx->blas;
    rocblas_status bs[5];
    if (cx) {
      const rocblas_double_complex one = {1.0, 0.0};
      auto* m = (rocblas_double_complex*)M;
      auto* a = (rocblas_double_complex*)Hs;
      bs[0] = factored ? rocblas_status_success : rocsolver_zpotrf(bh, rocblas_fill_upper, ns, m, ns, info);
      bs[1] = rocblas_ztrsm(bh, rocblas_side_left, rocblas_fill_upper, rocblas_operation_conjugate_transpose,
                            rocblas_diagonal_non_unit, ns, ns, &one, m, ns, a, ns);
      bs[2] = rocblas_ztrsm(bh, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, ns, ns, &one, m, ns, a, ns);
      if (own_trd) {
        // A = Q T Q^H (one workgroup), T = Z diag(D) Z^T, eigenvectors Q Z
        auto* x = (rocblas_double_complex*)X;
        if ((rc = reduce(a)) != CYX_OK) return rc;
        bs[3] = rocsolver_zstedc(bh, rocblas_evect_tridiagonal, ns, D, E, x, ns, info + 1);
        if (bs[3] == rocblas_status_success)
          bs[3] = rocsolver_zunmtr(bh, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, ns, ns, a, ns,
                                   (rocblas_double_complex*)tau, x, ns);
        a = x;
      } else {
        bs[3] = rocsolver_zheevd(bh, rocblas_evect_original, rocblas_fill_upper, ns, a, ns, D, E, info + 1);
      }
      bs[4] = rocblas_ztrsm(bh, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, ns, ns, &one, m, ns, a, ns);
      Q = a;
    } else {
      const double one = 1.0;
      auto* m = (double*)M;
      auto* a = (double*)Hs;
      bs[0] = factored ? rocblas_status_success : rocsolver_dpotrf(bh, rocblas_fill_upper, ns, m, ns, info);
      bs[1] = rocblas_dtrsm(bh, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                            rocblas_diagonal_non_unit, ns, ns, &one, m, ns, a, ns);
      bs[2] = rocblas_dtrsm(bh, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, ns, ns, &one, m, ns, a, ns);
      if (own_trd) {
        auto* x = (double*)X;
        if ((rc = reduce(a)) != CYX_OK) return rc;
        bs[3] = rocsolver_dstedc(bh, rocblas_evect_tridiagonal, ns, D, E, x, ns, info + 1);
        if (bs[3] == rocblas_status_success)
          bs[3] = rocsolver_dormtr(bh, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, ns, ns, a, ns,
                                   (double*)tau, x, ns);
        a = x;
      } else {
        bs[3] = rocsolver_dsyevd(bh, rocblas_evect_original, rocblas_fill_upper, ns, a, ns, D, E, info + 1);
      }
      bs[4] = rocblas_dtrsm(bh, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, ns, ns, &one, m, ns, a, ns);
      Q = a;
    }
    for (int k = 0; k < 5; ++k)
      if (bs[k] != rocblas_status_success) {
        set_error("dense Rayleigh-Ritz step %d failed with rocBLAS/rocSOLVER status %d", k, (int)bs[k]);
        return CYX_ERR_SOLVER;
      }
  }
  // the FP32 rotation's error grows like eps32 / sqrt(min_i R_ii^2 / M_ii):
  // it runs for the blocks the overlap certified above
  if (use32) {
    if ((rc = rotate32(h, dtype, psi, Q, other, ns, N)) != CYX_OK) return rc;
  } else {
    if ((rc = rotate(h, dtype, psi, Q, other, ns, N)) != CYX_OK) return rc;
  }
  if ((rc = wnormalize(h, dtype, other, ns, N, scal)) != CYX_OK) return rc;
  rocblas_int hinfo[2] = {0, 0};
  unsigned trd_failed = 0;
  CYX_HIP(hipMemcpyAsync(eigval, D, ns * sizeof(double), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipMemcpyAsync(hinfo, info, 2 * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  if (trd_flag) CYX_HIP(hipMemcpyAsync(&trd_failed, trd_flag, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  if (trd_failed != 0) {
    set_error("tridiagonal reduction: a workgroup barrier timed out (workgroups not co-resident?)");
    return CYX_ERR_SOLVER;
  }
  if (hinfo[0] != 0) {
    set_error("overlap matrix not positive definite (potrf info = %d): the filtered block is rank deficient",
              (int)hinfo[0]);
    return CYX_ERR_SOLVER;
  }
  if (hinfo[1] != 0) {
    set_error("projected eigensolver did not converge (heevd info = %d)", (int)hinfo[1]);
    return CYX_ERR_SOLVER;
  }
  *result = other;
  return CYX_OK;
}

// Chebyshev recurrence over buffers A (holds X) and B; returns the buffer holding Y
static int cheb_impl(cyx_ham_s* h, int dtype, void* A, void* B, int ncol, int64_t ld, int m, double a, double b,
                     double a0, void** result, const KbScratch* ks = nullptr) {
  CYX_CHECK(m >= 1, "Chebyshev degree must be >= 1");
  CYX_CHECK(b > a && a0 < a, "filter needs a0 < a < b (got a0=%g a=%g b=%g)", a0, a, b);
  const double e = (b - a) / 2;
  const double c = (b + a) / 2;
  double sigma = e / (a0 - c);
  const double sigma1 = sigma;
  const double gamma = 2 / sigma1;
  int rc;
  void* X = A;
  void* Y = B;
  if ((rc = stencil_step(h, dtype, X, nullptr, Y, ncol, ld, sigma1 / e, c, 0.0, ks)) != CYX_OK) return rc;
  for (int i = 2; i <= m; ++i) {
    const double sigma2 = 1 / (gamma - sigma);
    if ((rc = stencil_step(h, dtype, Y, X, X, ncol, ld, 2 * sigma2 / e, c, sigma * sigma2, ks)) != CYX_OK) return rc;
    std::swap(X, Y);
    sigma = sigma2;
  }
  *result = Y;
  return CYX_OK;
}

}  // namespace cyx

using namespace cyx;

static int resolve_marks(cyx_ctx_s* c);

extern "C" {

const char* cyx_last_error(void) { return g_err; }
const char* cyx_version(void) { return "cyclix_msparc_amd 0.1 (gfx950)"; }

int cyx_ctx_create(int device, cyx_ctx_t* out) {
  CYX_CHECK(out != nullptr, "out is null");
  int n = 0;
  CYX_HIP(hipGetDeviceCount(&n));
  CYX_CHECK(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
  CYX_HIP(hipSetDevice(device));
  // device address of the no-allocation scratch, once per device
  int rc0;
  if ((rc0 = noalloc_scratch_init(device)) != CYX_OK) return rc0;
  cyx_ctx_s* c = new cyx_ctx_s();
  c->device = device;
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
      c->ncu = ncu;
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    set_error("hipStreamCreate failed");
    return CYX_ERR_HIP;
  }
  c->own_stream = true;
  rocblas_handle bh = nullptr;
  if (rocblas_create_handle(&bh) != rocblas_status_success) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    set_error("rocblas_create_handle failed");
    return CYX_ERR_HIP;
  }
  c->blas = (void*)bh;
  rocblas_set_stream((rocblas_handle)c->blas, c->stream);
  *out = c;
  return CYX_OK;
}

int cyx_ctx_destroy(cyx_ctx_t c) {
  if (!c) return CYX_OK;
  (void)hipSetDevice(c->device);
  (void)resolve_marks(c);
  if (c->blas) rocblas_destroy_handle((rocblas_handle)c->blas);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CYX_OK;
}

int cyx_ctx_set_stream(cyx_ctx_t c, void* s) {
  CYX_CHECK(c != nullptr, "ctx is null");
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = (hipStream_t)s;
  c->own_stream = false;
  rocblas_set_stream((rocblas_handle)c->blas, c->stream);
  return CYX_OK;
}

static const char* kPhaseNames[PH_COUNT] = {"stencil", "h_apply",  "kb_gather", "gram",       "rotate",
                                            "wnorm",   "dense_eig", "aar",      "rr_apply", "sternheimer"};

static int resolve_marks(cyx_ctx_s* c) {
  if (c->marks.empty()) return CYX_OK;
  CYX_HIP(hipStreamSynchronize(c->stream));
  for (auto& m : c->marks) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
      c->phase_ms[m.phase] += ms;
      c->phase_count[m.phase] += 1;
    }
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  c->marks.clear();
  return CYX_OK;
}

int cyx_ctx_set_timing(cyx_ctx_t c, int enable) {
  CYX_CHECK(c != nullptr, "ctx is null");
  int rc = resolve_marks(c);
  c->timing = enable != 0;
  return rc;
}

int cyx_ctx_reset_timing(cyx_ctx_t c) {
  CYX_CHECK(c != nullptr, "ctx is null");
  int rc = resolve_marks(c);
  for (int k = 0; k < PH_COUNT; ++k) {
    c->phase_ms[k] = 0;
    c->phase_count[k] = 0;
  }
  return rc;
}

int cyx_ctx_read_timing(cyx_ctx_t c, const char* phase, double* total_ms, int64_t* count) {
  CYX_CHECK(c && phase && total_ms && count, "null argument");
  int rc = resolve_marks(c);
  if (rc != CYX_OK) return rc;
  for (int k = 0; k < PH_COUNT; ++k)
    if (std::strcmp(phase, kPhaseNames[k]) == 0) {
      *total_ms = c->phase_ms[k];
      *count = c->phase_count[k];
      return CYX_OK;
    }
  set_error("unknown phase '%s'", phase);
  return CYX_ERR_ARG;
}

int cyx_ctx_synchronize(cyx_ctx_t c) {
  CYX_CHECK(c != nullptr, "ctx is null");
  CYX_HIP(hipStreamSynchronize(c->stream));
  return CYX_OK;
}

int cyx_ham_create(cyx_ctx_t ctx, const cyx_grid_desc* grid, const double kpt[3], const double* veff,
                   const double* W, const cyx_projector_desc* proj, cyx_ham_t* out) {
  CYX_CHECK(ctx && out && veff && W && kpt, "null argument");
  int rc;
  if ((rc = check_grid(grid)) != CYX_OK) return rc;
  CYX_CHECK(kpt[0] == 0.0, "radial Bloch wavenumber must be 0 (Dirichlet r)");
  CYX_HIP(hipSetDevice(ctx->device));
  cyx_ham_s* h = new cyx_ham_s();
  h->ctx = ctx;
  h->device = ctx->device;
  h->g = *grid;
  for (int d = 0; d < 3; ++d) h->kpt[d] = kpt[d];
  h->bloch = kpt[1] != 0.0 || kpt[2] != 0.0;
  const int64_t N = h->N();
  const double L2 = grid->nt * grid->dt, L3 = grid->nz * grid->dz;
  // blochLaplacian_1d phases (src/blochLaplacian_1d.m:25-39): exactly 1 when the component is 0
  const double a2 = kpt[1] * L2, a3 = kpt[2] * L3;
  h->ph[0][0] = kpt[1] == 0 ? 1 : std::cos(-a2);
  h->ph[0][1] = kpt[1] == 0 ? 0 : std::sin(-a2);
  h->ph[1][0] = kpt[1] == 0 ? 1 : std::cos(a2);
  h->ph[1][1] = kpt[1] == 0 ? 0 : std::sin(a2);
  h->ph[2][0] = kpt[2] == 0 ? 1 : std::cos(-a3);
  h->ph[2][1] = kpt[2] == 0 ? 0 : std::sin(-a3);
  h->ph[3][0] = kpt[2] == 0 ? 1 : std::cos(a3);
  h->ph[3][1] = kpt[2] == 0 ? 0 : std::sin(a3);
  auto fail = [&](int code) {
    cyx_ham_destroy(h);
    return code;
  };
  if (hipMalloc((void**)&h->W, N * sizeof(double)) != hipSuccess) return fail(CYX_ERR_NOMEM);
  if (hipMemcpy(h->W, W, N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return fail(CYX_ERR_HIP);
  {
    // metric weights W' = r dr dt dz (no end corrections): H is self-adjoint in them
    std::vector<double> wm(N);
    for (int64_t n = 0; n < N; ++n) {
      const int i = (int)(n % grid->nr);
      wm[n] = (grid->xin + i * grid->dr) * grid->dr * grid->dt * grid->dz;
    }
    if (hipMalloc((void**)&h->Wm, N * sizeof(double)) != hipSuccess) return fail(CYX_ERR_NOMEM);
    if (hipMemcpy(h->Wm, wm.data(), N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return fail(CYX_ERR_HIP);
  }
  const std::vector<double> rcf = radial_coef(*grid, -0.5);
  if ((rc = upload_coef(h->c64, *grid, rcf, veff, N)) != CYX_OK) return fail(rc);
  if ((rc = upload_coef(h->c32, *grid, rcf, veff, N)) != CYX_OK) return fail(rc);
  if (proj && proj->n_atom > 0) {
    for (int im = 0; im < proj->n_image; ++im) {
      const int a = proj->image_atom[im];
      if (a < 0 || a >= proj->n_atom) {
        set_error("image %d has atom index %d out of range", im, a);
        return fail(CYX_ERR_ARG);
      }
    }
    int64_t off = 0;
    for (int im = 0; im < proj->n_image; ++im)
      for (int e = 0; e < proj->image_nnodes[im]; ++e, ++off)
        if (proj->rc_pos[off] < 0 || proj->rc_pos[off] >= N) {
          set_error("rc_pos[%lld] = %d outside the grid", (long long)off, proj->rc_pos[off]);
          return fail(CYX_ERR_ARG);
        }
    if (proj->n_image > 4096) {
      set_error("too many projector images (%d > 4096)", proj->n_image);
      return fail(CYX_ERR_ARG);
    }
    if ((rc = upload_kb(h->kb_z, proj, W, kpt, N)) != CYX_OK) return fail(rc);
    if ((rc = upload_kb(h->kb_c, proj, W, kpt, N)) != CYX_OK) return fail(rc);
    if (!h->bloch) {
      if ((rc = upload_kb(h->kb_d, proj, W, kpt, N)) != CYX_OK) return fail(rc);
      if ((rc = upload_kb(h->kb_f, proj, W, kpt, N)) != CYX_OK) return fail(rc);
    }
    h->nproj = h->kb_z.nproj;
  }
  *out = h;
  return CYX_OK;
}

int cyx_ham_set_veff(cyx_ham_t h, const double* veff, int on_device) {
  CYX_CHECK(h && veff, "null argument");
  const int64_t N = h->N();
  if (on_device) return veff_from_device(h, veff);
  CYX_HIP(hipStreamSynchronize(h->ctx->stream));
  CYX_HIP(hipMemcpy(h->c64.veff, veff, N * sizeof(double), hipMemcpyHostToDevice));
  std::vector<float> vf(veff, veff + N);
  CYX_HIP(hipMemcpy(h->c32.veff, vf.data(), N * sizeof(float), hipMemcpyHostToDevice));
  return CYX_OK;
}

int cyx_ham_destroy(cyx_ham_t h) {
  if (!h) return CYX_OK;
  (void)hipSetDevice(h->device);  // not through h->ctx: the context may already be gone
  (void)hipDeviceSynchronize();
  (void)hipFree(h->W);
  (void)hipFree(h->Wm);
  (void)hipFree(h->c64.rcoef);
  (void)hipFree(h->c64.veff);
  (void)hipFree(h->c32.rcoef);
  (void)hipFree(h->c32.veff);
  free_kb(h->kb_d);
  free_kb(h->kb_f);
  free_kb(h->kb_z);
  free_kb(h->kb_c);
  (void)hipFree(h->kbC.p);
  (void)hipFree(h->small.p);
  (void)hipFree(h->sub.p);
  (void)hipFree(h->red.p);
  (void)hipFree(h->stern.p);
  cyx::spectral_destroy(h->spec);
  delete h;
  return CYX_OK;
}

const double* cyx_ham_weights(cyx_ham_t h) { return h ? h->W : nullptr; }

int cyx_ham_rotation_counts(cyx_ham_t h, int64_t* fp32, int64_t* fp64) {
  CYX_CHECK(h && fp32 && fp64, "null argument");
  *fp32 = h->rot32_count;
  *fp64 = h->rot64_count;
  return CYX_OK;
}

int cyx_ham_set_tuning(cyx_ham_t h, const cyx_tuning* t) {
  CYX_CHECK(h && t, "null argument");
  CYX_CHECK(t->tr >= 0 && t->zb >= 0 && t->flags >= 0 && t->flags < 8, "bad tuning values");
  h->tune_tr = t->tr;
  h->tune_zb = t->zb;
  h->tune_flags = t->flags;
  return CYX_OK;
}

int cyx_h_apply(cyx_ham_t h, int dtype, const void* X, void* HX, int ncol, int64_t ld) {
  CYX_CHECK(h && X && HX, "null argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ncol >= 1 && ld >= h->N(), "bad block shape (ncol=%d ld=%lld N=%lld)", ncol, (long long)ld,
            (long long)h->N());
  CYX_CHECK(X != HX, "h_apply is out of place");
  return stencil_step(h, dtype, X, nullptr, HX, ncol, ld, 1.0, 0.0, 0.0);
}

int cyx_chebyshev_filter(cyx_ham_t h, int dtype, void* X, void* Y, int ncol, int64_t ld, int m, double a, double b,
                         double a0) {
  CYX_CHECK(h && X && Y && X != Y, "null or aliased argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ncol >= 1 && ld >= h->N(), "bad block shape");
  void* res = nullptr;
  if ((rc = cheb_impl(h, dtype, X, Y, ncol, ld, m, a, b, a0, &res)) != CYX_OK) return rc;
  if (res != Y)
    CYX_HIP(hipMemcpy2DAsync(Y, ld * dtype_size(dtype), res, ld * dtype_size(dtype), h->N() * dtype_size(dtype), ncol,
                             hipMemcpyDeviceToDevice, h->ctx->stream));
  return CYX_OK;
}

int cyx_chebyshev_filter_noalloc(cyx_ham_t h, int dtype, void* X, void* Y, int ncol, int64_t ld, int m, double a,
                                 double b, double a0) {
  CYX_CHECK(h && X && Y && X != Y, "null or aliased argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ncol >= 1 && ld >= h->N(), "bad block shape");
  // the scratch bound is checked before anything is enqueued
  const size_t need = kb_scratch_bytes(h, dtype, ncol);
  CYX_CHECK(need <= NOALLOC_SCRATCH_BYTES,
            "cyx_chebyshev_filter_noalloc: %d columns need %zu bytes of KB scratch, above the %zu-byte maximum; "
            "use fewer columns per call or cyx_chebyshev_filter", ncol, need, NOALLOC_SCRATCH_BYTES);
  char* base = (char*)noalloc_scratch(h->device);
  CYX_CHECK(base != nullptr, "no scratch address for device %d (no context was created on it)", h->device);
  // C region as kb_scratch_bytes lays it out: ncol x nproj of the block's
  // type, rounded up to 256 bytes, then the partials
  const size_t c_region = ((size_t)ncol * h->nproj * dtype_size(dtype) + 255) & ~(size_t)255;
  const KbScratch ks{base, base + c_region};
  void* res = nullptr;
  if ((rc = cheb_impl(h, dtype, X, Y, ncol, ld, m, a, b, a0, &res, &ks)) != CYX_OK) return rc;
  if (res != Y)
    CYX_HIP(hipMemcpy2DAsync(Y, ld * dtype_size(dtype), res, ld * dtype_size(dtype), h->N() * dtype_size(dtype), ncol,
                             hipMemcpyDeviceToDevice, h->ctx->stream));
  return CYX_OK;
}

int64_t cyx_eigsolver_workspace_bytes(cyx_ham_t h, int dtype, int ns) {
  if (!h) return -1;
  return h->N() * (int64_t)ns * (int64_t)dtype_size(dtype);
}

int cyx_rayleigh_ritz(cyx_ham_t h, int dtype, void* psi, void* work, int ns, double* eigval) {
  CYX_CHECK(h && psi && work && eigval && psi != work, "null or aliased argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ns >= 1 && ns <= h->N(), "bad state count %d", ns);
  void* res = nullptr;
  if ((rc = rr_impl(h, dtype, psi, work, ns, eigval, &res)) != CYX_OK) return rc;
  if (res != psi)
    CYX_HIP(hipMemcpyAsync(psi, res, h->N() * ns * dtype_size(dtype), hipMemcpyDeviceToDevice, h->ctx->stream));
  CYX_HIP(hipStreamSynchronize(h->ctx->stream));
  return CYX_OK;
}

int cyx_eigsolver_pass(cyx_ham_t h, int dtype, void* psi, void* work, int ns, int npl, double lambda_cutoff,
                       double bup, double a0, double* eigval) {
  CYX_CHECK(h && psi && work && eigval && psi != work, "null or aliased argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ns >= 1 && ns <= h->N(), "bad state count %d", ns);
  const int64_t N = h->N();
  void* filt = nullptr;
  if ((rc = cheb_impl(h, dtype, psi, work, ns, N, npl, lambda_cutoff, bup, a0, &filt)) != CYX_OK) return rc;
  void* other = (filt == psi) ? work : psi;
  void* res = nullptr;
  if ((rc = rr_impl(h, dtype, filt, other, ns, eigval, &res)) != CYX_OK) return rc;
  if (res != psi) CYX_HIP(hipMemcpyAsync(psi, res, N * ns * dtype_size(dtype), hipMemcpyDeviceToDevice, h->ctx->stream));
  CYX_HIP(hipStreamSynchronize(h->ctx->stream));
  return CYX_OK;
}

int cyx_ritz_residuals(cyx_ham_t h, int dtype, const void* psi, void* work, int ns, const double* eigval,
                       double* resid) {
  CYX_CHECK(h && psi && work && eigval && resid && psi != work, "null or aliased argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ns >= 1 && ns <= h->N(), "bad state count %d", ns);
  const int64_t N = h->N();
  if ((rc = ensure(h->sub, std::max(h->sub.bytes, 2 * (size_t)ns * sizeof(double)))) != CYX_OK) return rc;
  double* lam = (double*)h->sub.p;
  double* res = lam + ns;
  hipStream_t st = h->ctx->stream;
  CYX_HIP(hipMemcpyAsync(lam, eigval, ns * sizeof(double), hipMemcpyHostToDevice, st));
  if ((rc = stencil_apply_fp64(h, dtype, psi, work, ns, N)) != CYX_OK) return rc;
  if ((rc = ritz_residuals(h, dtype, psi, work, lam, ns, N, res)) != CYX_OK) return rc;
  CYX_HIP(hipMemcpyAsync(resid, res, ns * sizeof(double), hipMemcpyDeviceToHost, st));
  CYX_HIP(hipStreamSynchronize(st));
  return CYX_OK;
}

int cyx_weighted_gram(cyx_ham_t h, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                      void* M) {
  CYX_CHECK(h && A && B && M, "null argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(na >= 1 && nb >= 1 && lda >= h->N() && ldb >= h->N(), "bad block shapes (na=%d nb=%d)", na, nb);
  if ((rc = gram_rect(h, dtype, A, lda, na, B, ldb, nb, h->W, M, false)) != CYX_OK) return rc;
  CYX_HIP(hipStreamSynchronize(h->ctx->stream));
  return CYX_OK;
}

int cyx_lanczos_bounds(cyx_ham_t h, const double* v0, int maxit, double tol, double* lmin, double* lmax) {
  CYX_CHECK(h && v0 && lmin && lmax, "null argument");
  CYX_CHECK(maxit >= 5 && tol > 0, "bad Lanczos parameters");
  if (h->bloch) return lanczos_impl<cplx<double>>(h, CYX_C128, v0, maxit, tol, lmin, lmax);
  return lanczos_impl<double>(h, CYX_F64, v0, maxit, tol, lmin, lmax);
}

int cyx_density_accumulate(cyx_ham_t h, int dtype, const void* psi, int ns, const double* occ, double weight,
                           double* rho) {
  CYX_CHECK(h && psi && occ && rho, "null argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ns >= 1, "bad state count");
  if ((rc = ensure(h->sub, std::max(h->sub.bytes, (size_t)ns * sizeof(double)))) != CYX_OK) return rc;
  double* occ_dev = (double*)h->sub.p;
  CYX_HIP(hipMemcpyAsync(occ_dev, occ, ns * sizeof(double), hipMemcpyHostToDevice, h->ctx->stream));
  rc = density_accumulate(h, dtype, psi, ns, occ_dev, 2.0 * weight, rho);
  if (rc == CYX_OK) CYX_HIP(hipStreamSynchronize(h->ctx->stream));
  return rc;
}

int cyx_lda_pw(cyx_ctx_t ctx, int64_t N, const double* rho, const double* phi, double* vxc, double* veff,
               double rhotol) {
  CYX_CHECK(ctx && rho && vxc && N > 0, "null argument");
  CYX_CHECK(veff == nullptr || phi != nullptr, "veff needs phi");
  return lda_pw(ctx->stream, N, rho, phi, vxc, veff, rhotol);
}

int cyx_density_response(cyx_ham_t h, int dtype, const void* psi, const void* dpsi, int ns, const double* occ,
                         const double* dg, double weight, double* drho) {
  CYX_CHECK(h && psi && dpsi && occ && drho, "null argument");
  int rc;
  if ((rc = check_dtype(h, dtype)) != CYX_OK) return rc;
  CYX_CHECK(ns >= 1, "bad state count");
  if ((rc = ensure(h->sub, std::max(h->sub.bytes, 2 * (size_t)ns * sizeof(double)))) != CYX_OK) return rc;
  double* coef = (double*)h->sub.p;
  std::vector<double> hc(2 * (size_t)ns, 0.0);
  for (int s = 0; s < ns; ++s) {
    hc[s] = occ[s];
    if (dg) hc[ns + s] = dg[s];
  }
  CYX_HIP(hipMemcpyAsync(coef, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice, h->ctx->stream));
  rc = density_response(h, dtype, psi, dpsi, ns, coef, 2.0 * weight, drho);
  if (rc == CYX_OK) CYX_HIP(hipStreamSynchronize(h->ctx->stream));
  return rc;
}

int cyx_lda_pw_fxc(cyx_ctx_t ctx, int64_t N, const double* rho, const double* drho, const double* dphi, double* fxc,
                   double* dveff, double rhotol) {
  CYX_CHECK(ctx && rho && N > 0, "null argument");
  CYX_CHECK(fxc || dveff, "no output");
  CYX_CHECK(dveff == nullptr || (drho && dphi), "dveff needs drho and dphi");
  return lda_pw_fxc(ctx->stream, N, rho, drho, dphi, fxc, dveff, rhotol);
}

int cyx_energy_terms(cyx_ham_t h, const double* rho, const double* vxc, const double* phi, const double* b,
                     double rhotol, double out[3]) {
  CYX_CHECK(h && rho && vxc && phi && b && out, "null argument");
  const int nblk = 512;
  int rc;
  if ((rc = ensure(h->red, 3 * (size_t)nblk * sizeof(double))) != CYX_OK) return rc;
  return energy_terms(h->ctx->stream, h->N(), rho, vxc, phi, b, h->W, rhotol, (double*)h->red.p, nblk, out);
}

int cyx_poisson_create(cyx_ctx_t ctx, const cyx_grid_desc* grid, cyx_poisson_t* out) {
  CYX_CHECK(ctx && out, "null argument");
  int rc;
  if ((rc = check_grid(grid, false)) != CYX_OK) return rc;
  CYX_HIP(hipSetDevice(ctx->device));
  cyx_poisson_s* p = new cyx_poisson_s();
  p->ctx = ctx;
  p->g = *grid;
  const std::vector<double> c = radial_coef(*grid, 1.0);
  std::vector<double> ct(c.size());
  // radius-fastest copy for the AAR kernels, with the radial taps that fall
  // outside [0, nr) set to zero (the stencil is truncated there), so the
  // kernels apply every tap without a branch
  for (int i = 0; i < grid->nr; ++i)
    for (int k = 0; k < RCOEF; ++k) {
      double v = c[(size_t)i * RCOEF + k];
      if (k < FDN && i + k + 1 >= grid->nr) v = 0.0;               // c+[p], p = k + 1
      if (k >= FDN && k < 2 * FDN && i - (k - FDN + 1) < 0) v = 0.0;  // c-[p], p = k - FDN + 1
      ct[(size_t)k * grid->nr + i] = v;
    }
  if (hipMalloc((void**)&p->coef, c.size() * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&p->coefT, c.size() * sizeof(double)) != hipSuccess) {
    (void)hipFree(p->coef);
    delete p;
    set_error("allocation failed");
    return CYX_ERR_NOMEM;
  }
  CYX_HIP(hipMemcpy(p->coef, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
  CYX_HIP(hipMemcpy(p->coefT, ct.data(), ct.size() * sizeof(double), hipMemcpyHostToDevice));
  *out = p;
  return CYX_OK;
}

int cyx_poisson_apply(cyx_poisson_t p, const double* x, double* y) {
  CYX_CHECK(p && x && y && x != y, "null or aliased argument");
  return poisson_apply(p, x, y);
}

int cyx_poisson_set_preconditioner(cyx_poisson_t p, cyx_unet_t net) {
  CYX_CHECK(p, "null argument");
  if (net) {
    CYX_CHECK(unet_matches(net, p->g), "the U-Net was built for another grid");
    CYX_CHECK(unet_context(net) == p->ctx, "the U-Net and the Poisson handle must share a context");
  }
  p->net = net;
  return CYX_OK;
}

int cyx_poisson_destroy(cyx_poisson_t p) {
  if (!p) return CYX_OK;
  (void)hipDeviceSynchronize();
  cyx::spectral_destroy(p->spec);
  (void)hipFree(p->coef);
  (void)hipFree(p->coefT);
  (void)hipFree(p->scratch);
  (void)hipFree(p->tbuf.p);
  delete p;
  return CYX_OK;
}

int cyx_poisson_set_spectral(cyx_poisson_t ps, int on) {
  CYX_CHECK(ps, "null argument");
  CYX_HIP(hipSetDevice(ps->ctx->device));
  if (on && !ps->spec) {
    int rc = cyx::spectral_create(ps->g, &ps->spec);
    if (rc != CYX_OK) return rc;
  }
  ps->use_spectral = on != 0;
  return CYX_OK;
}

int cyx_poisson_spectral_apply(cyx_poisson_t ps, const double* r, double* z, int ncol, int64_t ld) {
  CYX_CHECK(ps && r && z, "null argument");
  CYX_CHECK(ncol >= 1 && ld >= ps->N(), "invalid block shape (ncol %d, ld %lld)", ncol, (long long)ld);
  CYX_HIP(hipSetDevice(ps->ctx->device));
  if (!ps->spec) {
    int rc = cyx::spectral_create(ps->g, &ps->spec);
    if (rc != CYX_OK) return rc;
  }
  return cyx::spectral_solve_real(ps->spec, r, z, ncol, ld, ps->ctx->stream);
}

int cyx_aar(cyx_poisson_t ps, const double* f, double* x, int use_guess, double tol, int max_iter, double omega,
            double beta, int m, int p, cyx_aar_info* info) {
  CYX_CHECK(ps && f && x && f != x, "null or aliased argument");
  return poisson_run(ps, f, x, use_guess, tol, max_iter, omega, beta, m, p, info);
}

int cyx_poisson_solve(cyx_poisson_t ps, const double* rho, const double* b, double* phi, int is_guess, double tol,
                      int max_iter, cyx_aar_info* info) {
  CYX_CHECK(ps && rho && b && phi, "null argument");
  DeviceBuf fbuf;
  int rc;
  if ((rc = ensure(fbuf, ps->N() * sizeof(double))) != CYX_OK) return rc;
  double* f = (double*)fbuf.p;
  rc = poisson_rhs(ps, rho, b, f);
  if (rc == CYX_OK) rc = poisson_run(ps, f, phi, is_guess, tol, max_iter, 0.6, 0.6, 7, 6, info);
  (void)hipStreamSynchronize(ps->ctx->stream);
  (void)hipFree(fbuf.p);
  return rc;
}

}  // extern "C"
