// Host side of the fused Cyclix stencil: launch parameters of one recurrence
// step, the KB gather / scatter kernels around it, and the batched filter
// plans (many channels per launch, captured in a HIP graph).  The stencil
// kernels themselves are in cyx_stencil_kernel.hpp.
#include <cstring>
#include <mutex>
#include <type_traits>

#include "cyx_internal.hpp"
#include "cyx_reduce.hpp"
#include "cyx_stencil_kernel.hpp"

namespace cyx {


// ---------------------------------------------------------------- KB gather
// C[s][p] = gamma_p * sum_e g[e] * Y[s][node[e]]   (src/h_nonlocal_vector_mult.m:30-38)
template <class T>
__global__ void __launch_bounds__(256) kb_gather_kernel(const T* __restrict__ Y, int64_t ld, const int* __restrict__ g_ptr,
                                                        const int* __restrict__ g_node, const T* __restrict__ g_coef,
                                                        const typename traits<T>::real* __restrict__ gamma, int nproj,
                                                        T* __restrict__ C) {
  __shared__ T red[256];
  const int p = blockIdx.x;
  const int s = blockIdx.y;
  const T* y = Y + (int64_t)s * ld;
  T acc = zero<T>();
  const int e1 = g_ptr[p + 1];
  for (int e = g_ptr[p] + threadIdx.x; e < e1; e += blockDim.x) acc = cfma(g_coef[e], y[g_node[e]], acc);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = add(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) C[(int64_t)s * nproj + p] = rmul(gamma[p], red[0]);
}

// KB scatter as its own pass, after a stencil step without it:
// Y[s][n] += alpha sum_e cph_e C[s][proj_e], cph_e = chi_e e^{-i k.d_img(e)},
// over the nodes with entries (the second half of
// src/h_nonlocal_vector_mult.m:24-32).  A thread takes one node and KSC
// columns, so the node's entries are read once for KSC columns; consecutive
// threads take consecutive nodes of one column; the block's KSC columns of C
// are staged in LDS.
constexpr int KSC = 8;

template <class T, class TS = T>
__device__ __forceinline__ void kb_scatter_body(TS* __restrict__ Y, int64_t ld, int ncol,
                                                const int* __restrict__ kb_nodes, int nkbn,
                                                const uint32_t* __restrict__ s_pk, const T* __restrict__ s_cph,
                                                const int* __restrict__ s_pj, const T* __restrict__ C, int nproj,
                                                typename traits<T>::real alpha, int bx, int by, T* Cs) {
  const int s0 = by * KSC;
  for (int e = threadIdx.x; e < KSC * nproj; e += 256) {
    const int c = e / nproj, p = e % nproj;
    Cs[e] = s0 + c < ncol ? C[(int64_t)(s0 + c) * nproj + p] : zero<T>();
  }
  __syncthreads();
  const int q = bx * 256 + threadIdx.x;
  if (q >= nkbn) return;
  const int node = kb_nodes[q];
  // the node's values are loaded first, so that their latency overlaps the
  // entry loads and sums
  T yv[KSC];
#pragma unroll
  for (int c = 0; c < KSC; ++c) yv[c] = s0 + c < ncol ? to_arith<T>(Y[(int64_t)(s0 + c) * ld + node]) : zero<T>();
  const uint32_t pk = s_pk[node];
  const int h0 = (int)(pk >> 6), cnt = (int)(pk & 63u);
  T acc[KSC];
#pragma unroll
  for (int c = 0; c < KSC; ++c) acc[c] = zero<T>();
  for (int e = 0; e < cnt; ++e) {
    const T cph = s_cph[h0 + e];
    const int p = s_pj[h0 + e];
#pragma unroll
    for (int c = 0; c < KSC; ++c) acc[c] = cfma(cph, Cs[c * nproj + p], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < KSC; ++c)
    if (s0 + c < ncol) Y[(int64_t)(s0 + c) * ld + node] = to_store<TS>(rfma(alpha, acc[c], yv[c]));
}

template <class T, class TS = T>
__global__ void __launch_bounds__(256) kb_scatter_kernel(TS* __restrict__ Y, int64_t ld, int ncol,
                                                         const int* __restrict__ kb_nodes, int nkbn,
                                                         const uint32_t* __restrict__ s_pk,
                                                         const T* __restrict__ s_cph, const int* __restrict__ s_pj,
                                                         const T* __restrict__ C, int nproj,
                                                         typename traits<T>::real alpha) {
  extern __shared__ __align__(16) unsigned char ksc_raw[];
  kb_scatter_body<T, TS>(Y, ld, ncol, kb_nodes, nkbn, s_pk, s_cph, s_pj, C, nproj, alpha, blockIdx.x, blockIdx.y,
                         reinterpret_cast<T*>(ksc_raw));  // Cs: [KSC][nproj]
}

// Two-stage gather for the single-channel step.  The node range is cut into
// chunks shared by all projectors; block (pair, column group) sums
// g[e] * Y_s[node[e]] over one projector's entries in one chunk for GCOL
// columns, so each entry's node index and coefficient are loaded once for
// GCOL columns, and reduces in FP64 (fixed order) into part[s][pair].  The
// pairs of one chunk run back to back on one XCD (the launch order comes
// from the host tables), so the Y values that several projectors' footprints
// share are read from HBM once and then from that XCD's L2.
// kb_gather_reduce_kernel sums each projector's pairs (numbered projector-
// major, chunk order) and applies gamma_p.
constexpr int GCOL = 8;

template <class T, class TS = T>
__device__ __forceinline__ void kb_gather_part_body(const TS* __restrict__ Y, int64_t ld, int ncol,
                                                    const int* __restrict__ blk_pair, const int* __restrict__ pr_lo,
                                                    const int* __restrict__ pr_hi, const int* __restrict__ g_node,
                                                    const T* __restrict__ g_coef, int npair,
                                                    typename acc_of<T>::type* __restrict__ part, int bx, int by,
                                                    double* red) {
  constexpr bool CX = traits<T>::is_complex;
  constexpr int K = CX ? 2 * GCOL : GCOL;
  const int q = blk_pair[bx];
  if (q < 0) return;  // padding of the XCD-ordered launch (whole block)
  const int s0 = by * GCOL;
  const int nc = min(GCOL, ncol - s0);
  const int c0 = pr_lo[q], c1 = pr_hi[q];
  T acc[GCOL];
#pragma unroll
  for (int c = 0; c < GCOL; ++c) acc[c] = zero<T>();
  const TS* y = Y + (int64_t)s0 * ld;
  // the next entry's node index and coefficient are loaded one step ahead
  int e = c0 + threadIdx.x;
  int node = e < c1 ? g_node[e] : 0;
  T gc = e < c1 ? g_coef[e] : zero<T>();
  for (; e < c1; e += 256) {
    const int en = e + 256;
    const int node_n = en < c1 ? g_node[en] : 0;
    const T gc_n = en < c1 ? g_coef[en] : zero<T>();
    T yv[GCOL];
#pragma unroll
    for (int c = 0; c < GCOL; ++c) yv[c] = c < nc ? to_arith<T>(y[(int64_t)c * ld + node]) : zero<T>();
#pragma unroll
    for (int c = 0; c < GCOL; ++c) acc[c] = cfma(gc, yv[c], acc[c]);
    node = node_n;
    gc = gc_n;
  }
  double v[K];
#pragma unroll
  for (int c = 0; c < GCOL; ++c) {
    if constexpr (CX) {
      v[2 * c] = (double)acc[c].x;
      v[2 * c + 1] = (double)acc[c].y;
    } else {
      v[c] = (double)acc[c];
    }
  }
  const double sum = block_sums<K, 256>(v, red);
  if ((int)threadIdx.x < K) {
    const int c = CX ? threadIdx.x / 2 : threadIdx.x;
    if (c < nc) {
      double* out = reinterpret_cast<double*>(part + (int64_t)(s0 + c) * npair + q);
      out[CX ? (threadIdx.x & 1) : 0] = sum;
    }
  }
}

template <class T, class TS = T>
__global__ void __launch_bounds__(256) kb_gather_part_kernel(const TS* __restrict__ Y, int64_t ld, int ncol,
                                                             const int* __restrict__ blk_pair,
                                                             const int* __restrict__ pr_lo,
                                                             const int* __restrict__ pr_hi,
                                                             const int* __restrict__ g_node,
                                                             const T* __restrict__ g_coef, int npair,
                                                             typename acc_of<T>::type* __restrict__ part) {
  constexpr int K = traits<T>::is_complex ? 2 * GCOL : GCOL;
  __shared__ double red[(256 / 64) * K];
  kb_gather_part_body<T, TS>(Y, ld, ncol, blk_pair, pr_lo, pr_hi, g_node, g_coef, npair, part, blockIdx.x, blockIdx.y,
                             red);
}

// one wavefront per (column, projector): lane k sums pairs k, k + 64, ...
// of the projector (contiguous, chunk order), then a fixed butterfly
template <class T>
__device__ __forceinline__ void kb_gather_reduce_body(const typename acc_of<T>::type* __restrict__ part, int ncol,
                                                      int nproj, int npair, const int* __restrict__ pp_ptr,
                                                      const typename traits<T>::real* __restrict__ gamma,
                                                      T* __restrict__ C, int bx) {
  using D = typename acc_of<T>::type;
  const int idx = bx * 4 + (threadIdx.x >> 6);  // s * nproj + p
  const int lane = threadIdx.x & 63;
  if (idx >= ncol * nproj) return;  // whole wavefronts
  const int s = idx / nproj, p = idx - s * nproj;
  const D* ps = part + (int64_t)s * npair + pp_ptr[p];
  const int n = pp_ptr[p + 1] - pp_ptr[p];
  D acc{};
  for (int k = lane; k < n; k += 64) acc = add(acc, ps[k]);
  if constexpr (traits<T>::is_complex) {
    const double re = wave_sum(acc.x), im = wave_sum(acc.y);
    if (lane == 0) C[idx] = narrow<T>(rmul((double)gamma[p], D{re, im}));
  } else {
    const double v = wave_sum(acc);
    if (lane == 0) C[idx] = narrow<T>(rmul((double)gamma[p], D(v)));
  }
}

template <class T>
__global__ void __launch_bounds__(256) kb_gather_reduce_kernel(const typename acc_of<T>::type* __restrict__ part,
                                                               int ncol, int nproj, int npair,
                                                               const int* __restrict__ pp_ptr,
                                                               const typename traits<T>::real* __restrict__ gamma,
                                                               T* __restrict__ C) {
  kb_gather_reduce_body<T>(part, ncol, nproj, npair, pp_ptr, gamma, C, blockIdx.x);
}

// Batched filter plans run the same three KB passes for all channels in one
// launch each: entry c owns the blocks [blk0, blk0 + nblk) of the launch and
// maps its local block to the single-channel kernel's (x, y).  Gather-part
// entries start on a multiple of 8 blocks, so a channel's XCD-ordered pair
// table keeps its XCD assignment.
template <class T>
struct KbPartEntry {
  const T* Y;
  int64_t ld;
  const int *blk_pair, *pr_lo, *pr_hi, *g_node;
  const T* g_coef;
  typename acc_of<T>::type* part;
  int ncol, npair, nbx, blk0, nblk;
};
template <class T>
struct KbReduceEntry {
  const typename acc_of<T>::type* part;
  const int* pp_ptr;
  const typename traits<T>::real* gamma;
  T* C;
  int ncol, nproj, npair, blk0, nblk;
};
template <class T>
struct KbScatterEntry {
  T* Y;
  int64_t ld;
  const int* kb_nodes;
  const uint32_t* s_pk;
  const T* s_cph;
  const int* s_pj;
  const T* C;
  typename traits<T>::real alpha;
  int ncol, nkbn, nproj, nbx, blk0, nblk;
};

template <class E>
__device__ __forceinline__ int batch_entry(const E* __restrict__ Es, int nch) {
  int c = 0;
  while (c + 1 < nch && Es[c + 1].blk0 <= (int)blockIdx.x) ++c;
  return c;
}

template <class T>
__global__ void __launch_bounds__(256) kb_gather_part_batch_kernel(const KbPartEntry<T>* __restrict__ Es, int nch) {
  constexpr int K = traits<T>::is_complex ? 2 * GCOL : GCOL;
  __shared__ double red[(256 / 64) * K];
  const KbPartEntry<T>& G = Es[batch_entry(Es, nch)];
  const int loc = blockIdx.x - G.blk0;
  if (loc >= G.nblk) return;  // padding to the next multiple of 8
  kb_gather_part_body<T, T>(G.Y, G.ld, G.ncol, G.blk_pair, G.pr_lo, G.pr_hi, G.g_node, G.g_coef, G.npair, G.part,
                            loc % G.nbx, loc / G.nbx, red);
}

template <class T>
__global__ void __launch_bounds__(256) kb_gather_reduce_batch_kernel(const KbReduceEntry<T>* __restrict__ Es, int nch) {
  const KbReduceEntry<T>& G = Es[batch_entry(Es, nch)];
  kb_gather_reduce_body<T>(G.part, G.ncol, G.nproj, G.npair, G.pp_ptr, G.gamma, G.C, blockIdx.x - G.blk0);
}

template <class T>
__global__ void __launch_bounds__(256) kb_scatter_batch_kernel(const KbScatterEntry<T>* __restrict__ Es, int nch) {
  extern __shared__ __align__(16) unsigned char ksc_raw[];
  const KbScatterEntry<T>& S = Es[batch_entry(Es, nch)];
  const int loc = blockIdx.x - S.blk0;
  kb_scatter_body<T, T>(S.Y, S.ld, S.ncol, S.kb_nodes, S.nkbn, S.s_pk, S.s_cph, S.s_pj, S.C, S.nproj, S.alpha,
                        loc % S.nbx, loc / S.nbx, reinterpret_cast<T*>(ksc_raw));
}

// ---------------------------------------------------------------- launch
template <class T>
static const KbTables<T>& kb_of(cyx_ham_s* h);
template <>
const KbTables<double>& kb_of<double>(cyx_ham_s* h) {
  return h->kb_d;
}
template <>
const KbTables<float>& kb_of<float>(cyx_ham_s* h) {
  return h->kb_f;
}
template <>
const KbTables<cplx<double>>& kb_of<cplx<double>>(cyx_ham_s* h) {
  return h->kb_z;
}
template <>
const KbTables<cplx<float>>& kb_of<cplx<float>>(cyx_ham_s* h) {
  return h->kb_c;
}

template <class R>
static const OpCoef<R>& coef_of(cyx_ham_s* h);
template <>
const OpCoef<double>& coef_of<double>(cyx_ham_s* h) {
  return h->c64;
}
template <>
const OpCoef<float>& coef_of<float>(cyx_ham_s* h) {
  return h->c32;
}

// r tile: largest tile with tr * nt <= maxthr, balanced over nr; at least
// 2*FDN (one right-halo load per thread) unless the theta ring is too wide
// for that, then at least STENCIL_TR_MIN (the narrow-tile variant)
static int choose_tr(int nr, int nt, int maxthr) {
  int maxtr = maxthr / nt;
  if (maxtr > nr) maxtr = nr;
  const int lo = maxthr / nt >= 2 * FDN ? 2 * FDN : STENCIL_TR_MIN;
  if (maxtr < lo) maxtr = lo;
  int ntiles = (nr + maxtr - 1) / maxtr;
  int tr = (nr + ntiles - 1) / ntiles;
  if (tr < lo) tr = lo;
  return tr;
}

template <class T, class TS = T>
struct StepLaunch {
  StepParams<T, TS> P;
  long nblk;
  int nthr;
  size_t shm;
  bool hel, kb;
};

// Parameters of one recurrence step of one channel.  tr_force > 0 sets the
// radial tile (batched launches use one tile for all channels).
template <class T, class TS = T>
static int make_step(cyx_ham_s* h, const void* in_, const void* prev_, void* out_, int ncol, int64_t ld, double alpha,
                     double shift, double beta, int tr_force, int flags, StepLaunch<T, TS>& Lc) {
  using R = typename traits<T>::real;
  const cyx_grid_desc& g = h->g;
  const OpCoef<R>& C = coef_of<R>(h);
  const KbTables<T>& K = kb_of<T>(h);
  StepParams<T, TS> P{};
  P.in = (const TS*)in_;
  P.prev = (const TS*)prev_;
  P.out = (TS*)out_;
  P.ld = ld;
  P.nr = g.nr;
  P.nt = g.nt;
  P.nz = g.nz;
  P.tr = choose_tr(g.nr, g.nt, stencil_max_threads<T>());
  if (h->tune_tr > 0 && stencil_max_threads<T>() / g.nt >= 2 * FDN)
    P.tr = std::max(2 * FDN, std::min(h->tune_tr, stencil_max_threads<T>() / g.nt));
  if (tr_force > 0) P.tr = tr_force;
  CYX_CHECK(P.tr >= STENCIL_TR_MIN && P.tr * g.nt <= stencil_max_threads<T>(), "radial tile %d does not fit (nt=%d)",
            P.tr, g.nt);
  P.ntiles_r = ceil_div(g.nr, P.tr);
  // z planes per workgroup: the whole column, or z-blocks when few columns
  // leave compute units idle.  A z-block of zb planes costs zb + 2 FDN
  // staged planes (its halo), and a launch takes ceil(workgroups / resident
  // workgroup slots) rounds; the block length that minimises
  // rounds x (zb + 2 FDN) is taken (ties: the shorter).  Measured at
  // 120x24x96, complex128 H psi: 17 columns 133 -> 75 us, 8 columns 97 ->
  // 49 us; 128 and 256 columns keep whole columns.
  int zb = g.nz;
  {
    const long slots = (long)std::max(h->ctx->ncu, 1) * stencil_min_blocks<T>();
    double best = 1e300;
    for (int cand = g.nz;; cand = (cand + 1) / 2) {
      const long nwg = (long)P.ntiles_r * ceil_div(g.nz, cand) * ncol;
      const double cost = (double)((nwg + slots - 1) / slots) * (cand + 2 * FDN);
      if (cost <= best) {
        best = cost;
        zb = cand;
      }
      if (cand <= 4) break;
    }
  }
  if (h->tune_zb > 0) zb = std::min(h->tune_zb, g.nz);
  P.zb = zb;
  P.nontemporal = (flags & 1) ? 1 : 0;
  P.tile_major = (flags & 4) ? 1 : 0;
  P.ncol = ncol;
  P.nblocks_z = ceil_div(g.nz, zb);
  P.alpha = (R)alpha;
  P.beta = (R)beta;
  P.shift = (R)shift;
  P.rcoef = C.rcoef;
  P.veff = C.veff;
  for (int p = 0; p <= FDN; ++p) {
    P.w2[p] = C.w2[p];
    P.w1[p] = C.w1[p];
  }
  P.s_r2 = C.s_r2;
  P.s_r1 = C.s_r1;
  P.s_z2 = C.s_z2;
  P.s_x = C.s_x;
  {
    // z and cross-term weights with their scale folded in, formed in FP64
    double w2d[FDN + 1], w1d[FDN + 1];
    const double tau = (g.cell_typ == 3) ? 0.0 : g.twist;
    const double sz2 = -0.5 / (g.dz * g.dz), sx = (-2.0 * tau / g.dt) * (-0.5 / g.dz);
    host_fd_weights(w2d, w1d);
    for (int d = 0; d <= FDN; ++d) {
      P.wz2[d] = (R)(sz2 * w2d[d]);
      P.wx1[d] = (R)(sx * w1d[d]);
    }
  }
  P.xin = (R)g.xin;
  P.dr = (R)g.dr;
  P.ph_tl = phase_of<T>(h->ph[0][0], h->ph[0][1]);
  P.ph_tr = phase_of<T>(h->ph[1][0], h->ph[1][1]);
  P.ph_zl = phase_of<T>(h->ph[2][0], h->ph[2][1]);
  P.ph_zr = phase_of<T>(h->ph[3][0], h->ph[3][1]);
  P.s_pk = K.s_pk;
  P.s_chi = K.s_chi;
  P.s_il = K.s_il;
  P.il_phase = K.il_phase;
  P.il_proj = K.il_proj;
  P.nimgl = K.nimgl;
  P.kbC = (const T*)h->kbC.p;
  P.nproj = K.nproj;
  // LDS row stride: halo, then padding so that the next row starts on the
  // next bank group (row pitch minus tile width a multiple of 256 B)
  P.trh = stencil_trh<T>(P.tr);
  const long nblk = (long)P.ntiles_r * P.nblocks_z * ncol;
  CYX_CHECK(nblk < (1L << 31), "grid too large");
  P.xcd_remap = (nblk % 8 == 0 && !(flags & 2)) ? 1 : 0;
  Lc.P = P;
  Lc.nblk = nblk;
  Lc.nthr = P.tr * g.nt;
  Lc.shm = (2 * stencil_pp<T>() * (size_t)P.trh * (g.nt + 2 * FDN) + (size_t)K.nimgl) * sizeof(T);
  Lc.hel = g.cell_typ == 4 || g.cell_typ == 5;
  Lc.kb = K.nproj > 0;
  return CYX_OK;
}


// bytes of the KB scratch of one step: C, rounded up to 256 bytes, then the
// gather partials
template <class T>
static size_t kb_scratch_need(cyx_ham_s* h, int ncol) {
  const KbTables<T>& K = kb_of<T>(h);
  if (K.nproj == 0) return 0;
  using D = typename acc_of<T>::type;
  const size_t c = ((size_t)ncol * K.nproj * sizeof(T) + 255) & ~(size_t)255;
  return c + (size_t)ncol * std::max(K.npair, 1) * sizeof(D);
}

size_t kb_scratch_bytes(cyx_ham_s* h, int dtype, int ncol) {
  switch (dtype) {
    case CYX_F64:
      return kb_scratch_need<double>(h, ncol);
    case CYX_F32:
      return kb_scratch_need<float>(h, ncol);
    case CYX_C128:
      return kb_scratch_need<cplx<double>>(h, ncol);
    case CYX_C64:
      return kb_scratch_need<cplx<float>>(h, ncol);
  }
  return 0;
}

// KB coefficients C of `in`; ks (optional) supplies C and the partials, else
// the handle's lazily grown buffers are used.  Returns C through *C_out.
template <class T, class TS = T>
static int gather_launch(cyx_ham_s* h, const void* in_, int ncol, int64_t ld, hipStream_t st, const KbScratch* ks,
                         T** C_out) {
  const KbTables<T>& K = kb_of<T>(h);
  *C_out = nullptr;
  if (K.nproj == 0) return CYX_OK;
  using D = typename acc_of<T>::type;
  const int ngrp = ceil_div(ncol, GCOL);
  T* C;
  D* part;
  if (ks) {
    C = (T*)ks->C;
    part = (D*)ks->part;
  } else {
    CYX_CHECK(ensure(h->kbC, (size_t)ncol * K.nproj * sizeof(T)) == CYX_OK, "kb scratch allocation failed");
    // the partials share the Gram kernels' scratch (stream-ordered uses;
    // filter plans hold C buffers of their own, so kbC may grow here)
    CYX_CHECK(ensure(h->small, (size_t)ncol * std::max(K.npair, 1) * sizeof(D)) == CYX_OK,
              "kb scratch allocation failed");
    C = (T*)h->kbC.p;
    part = (D*)h->small.p;
  }
  // every partial (column, pair) is written by one block before the reduce
  // reads it, and the reduce writes every C entry: nothing needs clearing
  if (K.npair > 0)
    hipLaunchKernelGGL((kb_gather_part_kernel<T, TS>), dim3(K.nblk_pair, ngrp), dim3(256), 0, st, (const TS*)in_, ld,
                       ncol, K.blk_pair, K.pr_lo, K.pr_hi, K.g_node, K.g_coef, K.npair, part);
  hipLaunchKernelGGL(kb_gather_reduce_kernel<T>, dim3(ceil_div(ncol * K.nproj, 4)), dim3(256), 0, st,
                     (const D*)part, ncol, K.nproj, K.npair, K.pp_ptr, K.gamma, C);
  CYX_HIP(hipGetLastError());
  *C_out = C;
  return CYX_OK;
}

// phase: timer of the stencil launch (-1: stencil / h_apply by whether prev is given)
template <class T, class TS = T>
static int step_impl(cyx_ham_s* h, const void* in_, const void* prev_, void* out_, int ncol, int64_t ld, double alpha,
                     double shift, double beta, int phase = -1, const KbScratch* ks = nullptr) {
  hipStream_t st = h->ctx->stream;
  int rc;
  T* C = nullptr;
  {
    PhaseTimer tm(h->ctx, PH_KB_GATHER);
    if ((rc = gather_launch<T, TS>(h, in_, ncol, ld, st, ks, &C)) != CYX_OK) return rc;
  }
  StepLaunch<T, TS> L;
  if ((rc = make_step<T, TS>(h, in_, prev_, out_, ncol, ld, alpha, shift, beta, 0, h->tune_flags, L)) != CYX_OK)
    return rc;
  L.P.kbC = C;
  // the KB scatter runs as its own pass after the stencil: the barrier-
  // synchronised plane loop then carries no per-node table loads
  const KbTables<T>& K = kb_of<T>(h);
  // (systems with so many projectors that the scatter's LDS slice of C would
  // not fit keep the scatter inside the stencil)
  const bool fits = (size_t)KSC * K.nproj * sizeof(T) <= 32 * 1024;
  // (the narrow-tile variant of wide theta rings has no KB term inside)
  CYX_CHECK(!L.kb || fits || L.P.tr >= 2 * FDN, "%d projectors on a theta ring of %d nodes are not supported", K.nproj,
            h->g.nt);
  const bool kb = L.kb && fits;
  if (kb) L.kb = false;
  {
    PhaseTimer tm(h->ctx, phase >= 0 ? phase : (prev_ ? PH_STENCIL : PH_HAPPLY));
    launch_step<T, TS>(L.P, (int)L.nblk, L.nthr, L.shm, L.hel, prev_ != nullptr, L.kb, st);
    CYX_HIP(hipGetLastError());
  }
  if (kb && K.nkbn > 0) {
    PhaseTimer tm(h->ctx, PH_KB_GATHER);
    hipLaunchKernelGGL((kb_scatter_kernel<T, TS>), dim3(ceil_div(K.nkbn, 256), ceil_div(ncol, KSC)), dim3(256),
                       (size_t)KSC * K.nproj * sizeof(T), st, (TS*)out_, ld, ncol, (const int*)K.kb_nodes, K.nkbn,
                       (const uint32_t*)K.s_pk, (const T*)K.s_cph, (const int*)K.s_pj, (const T*)C, K.nproj,
                       (typename traits<T>::real)alpha);
    CYX_HIP(hipGetLastError());
  }
  return CYX_OK;
}

int stencil_step(cyx_ham_s* h, int dtype, const void* in, const void* prev, void* out, int ncol, int64_t ld,
                 double alpha, double shift, double beta, const KbScratch* ks) {
  switch (dtype) {
    case CYX_F64:
      return step_impl<double>(h, in, prev, out, ncol, ld, alpha, shift, beta, -1, ks);
    case CYX_F32:
      return step_impl<float>(h, in, prev, out, ncol, ld, alpha, shift, beta, -1, ks);
    case CYX_C128:
      return step_impl<cplx<double>>(h, in, prev, out, ncol, ld, alpha, shift, beta, -1, ks);
    case CYX_C64:
      return step_impl<cplx<float>>(h, in, prev, out, ncol, ld, alpha, shift, beta, -1, ks);
  }
  set_error("unknown dtype %d", dtype);
  return CYX_ERR_ARG;
}

// ---------------------------------------------------------------- no-allocation scratch
// The device scratch of the no-allocation entry points (the KB coefficients
// and gather partials of their stencil steps).  Each GPU holds its own copy
// for the life of the process; every call on one GPU shares it, so calls
// that overlap on one GPU race.  In host code `scratch` names a host-side
// stand-in, never the device memory: its device address is looked up once
// per device and only that address is used.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch_init(int device) {
  CYX_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return CYX_OK;
  void* p = nullptr;
  CYX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CYX_CHECK(p != nullptr, "no device address for the scratch array on device %d", device);
  g_noalloc_ptr[device] = p;
  return CYX_OK;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  return g_noalloc_ptr[device];
}

int stencil_apply_fp64(cyx_ham_s* h, int dtype, const void* in, void* out, int ncol, int64_t ld) {
  switch (dtype) {
    case CYX_F64:
      return step_impl<double>(h, in, nullptr, out, ncol, ld, 1.0, 0.0, 0.0, PH_RR_APPLY);
    case CYX_F32:
      return step_impl<double, float>(h, in, nullptr, out, ncol, ld, 1.0, 0.0, 0.0, PH_RR_APPLY);
    case CYX_C128:
      return step_impl<cplx<double>>(h, in, nullptr, out, ncol, ld, 1.0, 0.0, 0.0, PH_RR_APPLY);
    case CYX_C64:
      return step_impl<cplx<double>, cplx<float>>(h, in, nullptr, out, ncol, ld, 1.0, 0.0, 0.0, PH_RR_APPLY);
  }
  set_error("unknown dtype %d", dtype);
  return CYX_ERR_ARG;
}

// ---------------------------------------------------------------- potential-only operator
// Y = alpha V X (FP64 arithmetic), the local part of potential_apply
template <class T>
__global__ void __launch_bounds__(256) pot_mul_kernel(const T* __restrict__ X, int64_t ldx, T* __restrict__ Y,
                                                      int64_t ldy, const double* __restrict__ V, int64_t N,
                                                      double alpha) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const int s = blockIdx.y;
  Y[(int64_t)s * ldy + n] = narrow<T>(rmul(alpha * V[n], widen(X[(int64_t)s * ldx + n])));
}

template <class T>
static int potential_impl(cyx_ham_s* h, const void* X, int64_t ldx, void* Y, int64_t ldy, int ncol, double alpha,
                          const KbScratch* ks) {
  hipStream_t st = h->ctx->stream;
  const int64_t N = h->N();
  const KbTables<T>& K = kb_of<T>(h);
  CYX_CHECK((size_t)KSC * K.nproj * sizeof(T) <= 32 * 1024, "too many projectors for the KB scatter (%d)", K.nproj);
  T* C = nullptr;
  int rc;
  {
    PhaseTimer tm(h->ctx, PH_KB_GATHER);
    if ((rc = gather_launch<T>(h, X, ncol, ldx, st, ks, &C)) != CYX_OK) return rc;
  }
  PhaseTimer tm(h->ctx, PH_HAPPLY);
  hipLaunchKernelGGL(pot_mul_kernel<T>, dim3(ceil_div(N, 256), ncol), dim3(256), 0, st, (const T*)X, ldx, (T*)Y, ldy,
                     (const double*)h->c64.veff, N, alpha);
  CYX_HIP(hipGetLastError());
  if (K.nproj > 0 && K.nkbn > 0) {
    hipLaunchKernelGGL((kb_scatter_kernel<T, T>), dim3(ceil_div(K.nkbn, 256), ceil_div(ncol, KSC)), dim3(256),
                       (size_t)KSC * K.nproj * sizeof(T), st, (T*)Y, ldy, ncol, (const int*)K.kb_nodes, K.nkbn,
                       (const uint32_t*)K.s_pk, (const T*)K.s_cph, (const int*)K.s_pj, (const T*)C, K.nproj,
                       (typename traits<T>::real)alpha);
    CYX_HIP(hipGetLastError());
  }
  return CYX_OK;
}

int potential_apply(cyx_ham_s* h, int dtype, const void* X, int64_t ldx, void* Y, int64_t ldy, int ncol,
                    double alpha, const KbScratch* ks) {
  switch (dtype) {
    case CYX_F64:
      return potential_impl<double>(h, X, ldx, Y, ldy, ncol, alpha, ks);
    case CYX_F32:
      return potential_impl<float>(h, X, ldx, Y, ldy, ncol, alpha, ks);
    case CYX_C128:
      return potential_impl<cplx<double>>(h, X, ldx, Y, ldy, ncol, alpha, ks);
    case CYX_C64:
      return potential_impl<cplx<float>>(h, X, ldx, Y, ldy, ncol, alpha, ks);
  }
  set_error("unknown dtype %d", dtype);
  return CYX_ERR_ARG;
}

}  // namespace cyx

// ---------------------------------------------------------------- batched filter plans

using namespace cyx;

struct cyx_filter_plan_s {
  cyx_ctx_s* ctx = nullptr;
  int dtype = 0, nch = 0, m = 0;
  std::vector<cyx_ham_s*> hams;
  std::vector<void*> A, B;
  std::vector<int> ncol;
  std::vector<int64_t> ld;
  int tr = 0, flags = 0;
  bool hel = false, kb = false;
  int nblk = 0, nthr = 0;
  size_t shm = 0;
  // per-channel KB coefficients C (ncol x nproj) and gather partials, owned
  // by the plan: the captured graph holds these addresses for the plan's
  // whole life, so they must not be the handle's lazily grown scratch
  std::vector<void*> kbC, kbPart;
  void* d_steps = nullptr;   // [m][nch] BatchEntry<T>, then [m][nch] KbScatterEntry<T>
  void* d_gather = nullptr;  // [2][nch] KbPartEntry<T> (inputs A, B), then [nch] KbReduceEntry<T>
  size_t step_bytes = 0, gather_bytes = 0, scat_off = 0, red_off = 0;
  int pblk = 0, rblk = 0, sblk = 0;  // blocks of the KB launches
  size_t sshm = 0;                   // scatter LDS: KSC x max nproj
  std::vector<unsigned char> h_steps;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;  // capture stream (the context's stream may be the null stream)
  // attached Sternheimer iterations: a second branch of the graph, captured
  // on cap2 between a fork and a join event
  cyx_stern_plan_s* stern = nullptr;
  int stern_iters = 0;
  hipStream_t cap2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

template <class T>
int plan_fill_steps(cyx_filter_plan_s* pl, const double* a, const double* b, const double* a0) {
  const int nch = pl->nch, m = pl->m;
  pl->h_steps.assign(pl->step_bytes, 0);
  BatchEntry<T>* E = reinterpret_cast<BatchEntry<T>*>(pl->h_steps.data());
  KbScatterEntry<T>* SE = reinterpret_cast<KbScatterEntry<T>*>(pl->h_steps.data() + pl->scat_off);
  for (int c = 0; c < nch; ++c) {
    CYX_CHECK(b[c] > a[c] && a0[c] < a[c], "channel %d: filter needs a0 < a < b (a0=%g a=%g b=%g)", c, a0[c], a[c],
              b[c]);
    // Chebyshev scalars of src/chebyshev_filter.m:24-42
    const double e = (b[c] - a[c]) / 2, cc = (b[c] + a[c]) / 2;
    double sigma = e / (a0[c] - cc);
    const double gamma = 2 / sigma;
    void* X = pl->A[c];
    void* Y = pl->B[c];
    int blk0 = 0;
    for (int i = 1; i <= m; ++i) {
      StepLaunch<T> L;
      int rc;
      if (i == 1) {
        rc = make_step<T>(pl->hams[c], X, nullptr, Y, pl->ncol[c], pl->ld[c], sigma / e, cc, 0.0, pl->tr, pl->flags, L);
      } else {
        const double sigma2 = 1 / (gamma - sigma);
        rc = make_step<T>(pl->hams[c], Y, X, X, pl->ncol[c], pl->ld[c], 2 * sigma2 / e, cc, sigma * sigma2, pl->tr,
                          pl->flags, L);
        std::swap(X, Y);
        sigma = sigma2;
      }
      if (rc != CYX_OK) return rc;
      L.P.kbC = nullptr;  // the KB term is the plan's own gather and scatter launches
      L.P.xcd_remap = 0;  // blocks of a channel do not start on an XCD boundary in general
      if (pl->kb) {
        // scatter of this step's output: Y_new += alpha KB C (C gathered from the step's input)
        const KbTables<T>& K = kb_of<T>(pl->hams[c]);
        KbScatterEntry<T>& se = SE[(size_t)(i - 1) * nch + c];
        se.Y = (T*)L.P.out;
        se.ld = pl->ld[c];
        se.kb_nodes = K.kb_nodes;
        se.s_pk = K.s_pk;
        se.s_cph = K.s_cph;
        se.s_pj = K.s_pj;
        se.C = (const T*)pl->kbC[c];
        se.alpha = L.P.alpha;
        se.ncol = pl->ncol[c];
        se.nkbn = K.nkbn;
        se.nproj = K.nproj;
        se.nbx = std::max(ceil_div(K.nkbn, 256), 1);
        se.nblk = se.nbx * ceil_div(pl->ncol[c], KSC);
        // channel offsets (same for every step)
        se.blk0 = c == 0 ? 0 : SE[(size_t)(i - 1) * nch + c - 1].blk0 + SE[(size_t)(i - 1) * nch + c - 1].nblk;
        if (i == 1 && c == nch - 1) pl->sblk = se.blk0 + se.nblk;
      }
      BatchEntry<T>& be = E[(size_t)(i - 1) * nch + c];
      be.P = L.P;
      be.nblk = (int)L.nblk;
      // channel offsets (same for every step)
      if (c == 0) be.blk0 = 0;
      else be.blk0 = E[(size_t)(i - 1) * nch + c - 1].blk0 + E[(size_t)(i - 1) * nch + c - 1].nblk;
      blk0 = be.blk0 + be.nblk;
      if (i == 1 && c == nch - 1) pl->nblk = blk0;
    }
  }
  return CYX_OK;
}

template <class T>
int plan_setup(cyx_filter_plan_s* pl) {
  using D = typename acc_of<T>::type;
  const int nch = pl->nch;
  int nt = pl->hams[0]->g.nt;
  pl->hel = false;
  pl->kb = false;
  size_t shm = 0;
  int nproj_max = 0;
  std::vector<KbPartEntry<T>> GP(2 * (size_t)nch);
  std::vector<KbReduceEntry<T>> GR((size_t)nch);
  int pblk = 0, rblk = 0;
  CYX_CHECK(pl->tr >= 2 * FDN, "batched plans take theta rings of at most %d nodes (got %d)",
            stencil_max_threads<T>() / (2 * FDN), nt);
  for (int c = 0; c < nch; ++c) {
    cyx_ham_s* h = pl->hams[c];
    CYX_CHECK(h->g.nt == nt, "batched channels must share nt (%d vs %d)", h->g.nt, nt);
    const bool hel = h->g.cell_typ == 4 || h->g.cell_typ == 5;
    const KbTables<T>& K = kb_of<T>(h);
    if (c == 0) {
      pl->hel = hel;
      pl->kb = K.nproj > 0;
    }
    CYX_CHECK(hel == pl->hel && (K.nproj > 0) == pl->kb,
              "batched channels must all be helical or all cyclic, and all with or without projectors");
    if (pl->kb && pl->kbC[c] == nullptr) {
      CYX_HIP(hipMalloc(&pl->kbC[c], (size_t)pl->ncol[c] * K.nproj * sizeof(T)));
      CYX_HIP(hipMalloc(&pl->kbPart[c], (size_t)pl->ncol[c] * std::max(K.npair, 1) * sizeof(D)));
    }
    nproj_max = std::max(nproj_max, K.nproj);
    const size_t trh = (size_t)stencil_trh<T>(pl->tr);
    shm = std::max(shm, 2 * stencil_pp<T>() * trh * (nt + 2 * FDN) * sizeof(T));
    if (!pl->kb) continue;
    // two-stage gather: the channel's XCD-ordered pair blocks x column groups,
    // starting on a multiple of 8 blocks
    const int nbx = K.nblk_pair, ngrp = ceil_div(pl->ncol[c], GCOL);
    for (int v = 0; v < 2; ++v) {
      KbPartEntry<T>& g = GP[(size_t)v * nch + c];
      g.Y = (const T*)(v == 0 ? pl->A[c] : pl->B[c]);
      g.ld = pl->ld[c];
      g.blk_pair = K.blk_pair;
      g.pr_lo = K.pr_lo;
      g.pr_hi = K.pr_hi;
      g.g_node = K.g_node;
      g.g_coef = K.g_coef;
      g.part = (D*)pl->kbPart[c];
      g.ncol = pl->ncol[c];
      g.npair = K.npair;
      g.nbx = std::max(nbx, 1);
      g.blk0 = pblk;
      g.nblk = K.npair > 0 ? nbx * ngrp : 0;
    }
    pblk += (GP[c].nblk + 7) & ~7;
    KbReduceEntry<T>& r = GR[c];
    r.part = (const D*)pl->kbPart[c];
    r.pp_ptr = K.pp_ptr;
    r.gamma = K.gamma;
    r.C = (T*)pl->kbC[c];
    r.ncol = pl->ncol[c];
    r.nproj = K.nproj;
    r.npair = K.npair;
    r.blk0 = rblk;
    r.nblk = ceil_div(pl->ncol[c] * K.nproj, 4);
    rblk += r.nblk;
  }
  pl->pblk = pblk;
  pl->rblk = rblk;
  pl->sshm = (size_t)KSC * nproj_max * sizeof(T);
  CYX_CHECK(!pl->kb || pl->sshm <= 64 * 1024, "too many projectors for the batched KB scatter (%d)", nproj_max);
  pl->shm = shm;
  pl->nthr = pl->tr * nt;
  pl->scat_off = ((size_t)pl->m * nch * sizeof(BatchEntry<T>) + 255) & ~(size_t)255;
  pl->step_bytes = pl->scat_off + (pl->kb ? (size_t)pl->m * nch * sizeof(KbScatterEntry<T>) : 0);
  pl->red_off = (2 * (size_t)nch * sizeof(KbPartEntry<T>) + 255) & ~(size_t)255;
  pl->gather_bytes = pl->red_off + (size_t)nch * sizeof(KbReduceEntry<T>);
  if (!pl->d_steps) {
    CYX_HIP(hipMalloc(&pl->d_steps, pl->step_bytes));
    CYX_HIP(hipMalloc(&pl->d_gather, pl->gather_bytes));
  }
  std::vector<unsigned char> hg(pl->gather_bytes, 0);
  std::memcpy(hg.data(), GP.data(), GP.size() * sizeof(KbPartEntry<T>));
  std::memcpy(hg.data() + pl->red_off, GR.data(), GR.size() * sizeof(KbReduceEntry<T>));
  CYX_HIP(hipMemcpy(pl->d_gather, hg.data(), pl->gather_bytes, hipMemcpyHostToDevice));
  return CYX_OK;
}

// the m steps of one filter pass, enqueued on st: per step the two-stage KB
// gather of the step's input, the stencil step without the KB term, and the
// KB scatter into its output, each one launch over all channels
template <class T>
void plan_enqueue(cyx_filter_plan_s* pl, hipStream_t st, int steps) {
  const BatchEntry<T>* E = (const BatchEntry<T>*)pl->d_steps;
  const KbScatterEntry<T>* SE = (const KbScatterEntry<T>*)((const unsigned char*)pl->d_steps + pl->scat_off);
  const KbPartEntry<T>* GP = (const KbPartEntry<T>*)pl->d_gather;
  const KbReduceEntry<T>* GR = (const KbReduceEntry<T>*)((const unsigned char*)pl->d_gather + pl->red_off);
  const int nch = pl->nch;
  // stencil input of step i: A for odd i, B for even i
  for (int i = 1; i <= steps; ++i) {
    if (pl->kb) {
      if (pl->pblk > 0)
        hipLaunchKernelGGL(kb_gather_part_batch_kernel<T>, dim3(pl->pblk), dim3(256), 0, st,
                           GP + ((i & 1) ? 0 : nch), nch);
      hipLaunchKernelGGL(kb_gather_reduce_batch_kernel<T>, dim3(pl->rblk), dim3(256), 0, st, GR, nch);
    }
    launch_batch<T>(E + (size_t)(i - 1) * nch, nch, pl->nblk, pl->nthr, pl->shm, pl->hel, i > 1, false,
                    stencil_nt_variant(pl->hams[0]->g.nt, pl->tr), st);
    if (pl->kb && pl->sblk > 0)
      hipLaunchKernelGGL(kb_scatter_batch_kernel<T>, dim3(pl->sblk), dim3(256), pl->sshm, st,
                         SE + (size_t)(i - 1) * nch, nch);
  }
}

template <class T>
int plan_run(cyx_filter_plan_s* pl, const double* a, const double* b, const double* a0) {
  int rc;
  if ((rc = plan_fill_steps<T>(pl, a, b, a0)) != CYX_OK) return rc;
  hipStream_t st = pl->ctx->stream;
  CYX_HIP(hipMemcpyAsync(pl->d_steps, pl->h_steps.data(), pl->step_bytes, hipMemcpyHostToDevice, st));
  if (!pl->exec) {
    // capture the pass once, on a stream of the plan's own (capture is not
    // allowed on the null stream); later passes replay it with new scalars
    if (!pl->cap) CYX_HIP(hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking));
    if (pl->stern && !pl->cap2) {
      CYX_HIP(hipStreamCreateWithFlags(&pl->cap2, hipStreamNonBlocking));
      CYX_HIP(hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming));
      CYX_HIP(hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming));
    }
    CYX_HIP(hipStreamBeginCapture(pl->cap, hipStreamCaptureModeThreadLocal));
    int rc2 = CYX_OK;
    hipError_t e2 = hipSuccess;
    if (pl->stern) {
      // the Sternheimer branch forks from the graph's start and joins at its
      // end; it shares no buffer with the filter steps
      e2 = hipEventRecord(pl->ev_fork, pl->cap);
      if (e2 == hipSuccess) e2 = hipStreamWaitEvent(pl->cap2, pl->ev_fork, 0);
      if (e2 == hipSuccess) rc2 = stern_plan_enqueue(pl->stern, pl->cap2, pl->stern_iters);
      if (e2 == hipSuccess) e2 = hipEventRecord(pl->ev_join, pl->cap2);
    }
    plan_enqueue<T>(pl, pl->cap, pl->m);
    if (pl->stern && e2 == hipSuccess) e2 = hipStreamWaitEvent(pl->cap, pl->ev_join, 0);
    hipError_t e = hipStreamEndCapture(pl->cap, &pl->graph);
    if (e == hipSuccess) e = e2;
    if (rc2 != CYX_OK || e != hipSuccess) {
      if (pl->graph) (void)hipGraphDestroy(pl->graph);
      pl->graph = nullptr;
      if (rc2 != CYX_OK) return rc2;
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return CYX_ERR_HIP;
    }
    CYX_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  }
  PhaseTimer tm(pl->ctx, PH_STENCIL);
  CYX_HIP(hipGraphLaunch(pl->exec, st));
  return CYX_OK;
}

// time one first-order step (reads A, writes B) per candidate (tile, flags)
template <class T>
int plan_autotune(cyx_filter_plan_s* pl, const double* a, const double* b, const double* a0) {
  const int nt = pl->hams[0]->g.nt;
  std::vector<int> trs;
  for (int tr : {20, 24, 30, 16, 21, 15, 12})
    if (tr >= 2 * FDN && tr * nt <= stencil_max_threads<T>() &&
        std::find(trs.begin(), trs.end(), tr) == trs.end())
      trs.push_back(tr);
  hipStream_t st = pl->ctx->stream;
  hipEvent_t e0, e1;
  CYX_HIP(hipEventCreate(&e0));
  CYX_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_tr = pl->tr, best_fl = pl->flags;
  const int saved_m = pl->m;
  pl->m = 1;
  int rc = CYX_OK;
  for (int tr : trs)
    for (int fl : {0, 1, 4}) {
      pl->tr = tr;
      pl->flags = fl;
      if ((rc = plan_setup<T>(pl)) != CYX_OK) break;
      if ((rc = plan_fill_steps<T>(pl, a, b, a0)) != CYX_OK) break;
      CYX_HIP(hipMemcpyAsync(pl->d_steps, pl->h_steps.data(), pl->step_bytes, hipMemcpyHostToDevice, st));
      plan_enqueue<T>(pl, st, 1);
      CYX_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 3; ++r) plan_enqueue<T>(pl, st, 1);
      CYX_HIP(hipEventRecord(e1, st));
      CYX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      CYX_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_tr = tr;
        best_fl = fl;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pl->m = saved_m;
  pl->tr = best_tr;
  pl->flags = best_fl;
  if (rc != CYX_OK) return rc;
  hipFree(pl->d_steps);
  pl->d_steps = nullptr;
  hipFree(pl->d_gather);
  pl->d_gather = nullptr;
  return plan_setup<T>(pl);
}

template <class T>
int plan_create(cyx_filter_plan_s* pl, int autotune, const double* a, const double* b, const double* a0) {
  pl->tr = choose_tr(pl->hams[0]->g.nr, pl->hams[0]->g.nt, stencil_max_threads<T>());
  pl->flags = 0;
  int rc;
  if (autotune) {
    const int saved_m = pl->m;
    pl->m = 1;
    rc = plan_setup<T>(pl);
    pl->m = saved_m;
    if (rc != CYX_OK) return rc;
    // the d_steps of the single-step setup is replaced after tuning
    if ((rc = plan_autotune<T>(pl, a, b, a0)) != CYX_OK) return rc;
    return CYX_OK;
  }
  return plan_setup<T>(pl);
}

}  // namespace

extern "C" {

int cyx_filter_plan_create(cyx_ctx_t ctx, int nch, const cyx_ham_t* hams, int dtype, void* const* A, void* const* B,
                           const int* ncol, const int64_t* ld, int m, int autotune, const double* a, const double* b,
                           const double* a0, cyx_filter_plan_t* out) {
  CYX_CHECK(ctx && hams && A && B && ncol && ld && out && nch >= 1 && m >= 1, "bad argument");
  CYX_CHECK(!autotune || (a && b && a0), "autotuning needs filter bounds");
  CYX_CHECK(dtype == CYX_F64 || dtype == CYX_C128 || dtype == CYX_F32 || dtype == CYX_C64, "bad dtype");
  cyx_filter_plan_s* pl = new cyx_filter_plan_s();
  pl->ctx = ctx;
  pl->dtype = dtype;
  pl->nch = nch;
  pl->m = m;
  for (int c = 0; c < nch; ++c) {
    if (!hams[c] || !A[c] || !B[c] || A[c] == B[c] || ncol[c] < 1 || ld[c] < hams[c]->N() || hams[c]->ctx != ctx ||
        (hams[c]->bloch && (dtype == CYX_F64 || dtype == CYX_F32))) {
      delete pl;
      set_error("channel %d: bad handle, buffers or shape", c);
      return CYX_ERR_ARG;
    }
    for (int d = 0; d < c; ++d)
      if (A[d] == A[c] || A[d] == B[c] || B[d] == A[c] || B[d] == B[c]) {
        delete pl;
        set_error("channels %d and %d share a buffer", d, c);
        return CYX_ERR_ARG;
      }
    pl->hams.push_back(hams[c]);
    pl->kbC.push_back(nullptr);
    pl->kbPart.push_back(nullptr);
    pl->A.push_back(A[c]);
    pl->B.push_back(B[c]);
    pl->ncol.push_back(ncol[c]);
    pl->ld.push_back(ld[c]);
  }
  int rc;
  switch (dtype) {
    case CYX_F64: rc = plan_create<double>(pl, autotune, a, b, a0); break;
    case CYX_F32: rc = plan_create<float>(pl, autotune, a, b, a0); break;
    case CYX_C128: rc = plan_create<cplx<double>>(pl, autotune, a, b, a0); break;
    default: rc = plan_create<cplx<float>>(pl, autotune, a, b, a0); break;
  }
  if (rc != CYX_OK) {
    (void)hipFree(pl->d_steps);
    (void)hipFree(pl->d_gather);
    for (void* p : pl->kbC) (void)hipFree(p);
    for (void* p : pl->kbPart) (void)hipFree(p);
    delete pl;
    return rc;
  }
  *out = pl;
  return CYX_OK;
}

int cyx_filter_plan_run(cyx_filter_plan_t pl, const double* a, const double* b, const double* a0, int* result_in_b) {
  CYX_CHECK(pl && a && b && a0, "null argument");
  int rc;
  switch (pl->dtype) {
    case CYX_F64: rc = plan_run<double>(pl, a, b, a0); break;
    case CYX_F32: rc = plan_run<float>(pl, a, b, a0); break;
    case CYX_C128: rc = plan_run<cplx<double>>(pl, a, b, a0); break;
    default: rc = plan_run<cplx<float>>(pl, a, b, a0); break;
  }
  if (result_in_b) *result_in_b = (pl->m % 2) ? 1 : 0;
  return rc;
}

int cyx_filter_plan_tuning(cyx_filter_plan_t pl, cyx_tuning* t) {
  CYX_CHECK(pl && t, "null argument");
  t->tr = pl->tr;
  t->zb = 0;
  t->flags = pl->flags;
  t->reserved = 0;
  return CYX_OK;
}

int cyx_filter_plan_attach_sternheimer(cyx_filter_plan_t pl, cyx_stern_plan_t sp, int iters) {
  CYX_CHECK(pl && (!sp || iters >= 1), "bad argument");
  CYX_CHECK(!sp || stern_plan_context(sp) == pl->ctx, "the Sternheimer plan belongs to another context");
  CYX_CHECK(!sp || !stern_plan_spectral(sp),
            "a Sternheimer plan with the spectral preconditioner runs its transforms outside a graph");
  // the next run captures the graph again
  if (pl->exec) {
    CYX_HIP(hipStreamSynchronize(pl->ctx->stream));
    (void)hipGraphExecDestroy(pl->exec);
  }
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  pl->exec = nullptr;
  pl->graph = nullptr;
  pl->stern = sp;
  pl->stern_iters = sp ? iters : 0;
  return CYX_OK;
}

int cyx_filter_plan_destroy(cyx_filter_plan_t pl) {
  if (!pl) return CYX_OK;
  (void)hipDeviceSynchronize();  // the context may already be gone
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  if (pl->cap) (void)hipStreamDestroy(pl->cap);
  if (pl->cap2) (void)hipStreamDestroy(pl->cap2);
  if (pl->ev_fork) (void)hipEventDestroy(pl->ev_fork);
  if (pl->ev_join) (void)hipEventDestroy(pl->ev_join);
  (void)hipFree(pl->d_steps);
  (void)hipFree(pl->d_gather);
  for (void* p : pl->kbC) (void)hipFree(p);
  for (void* p : pl->kbPart) (void)hipFree(p);
  delete pl;
  return CYX_OK;
}

}  // extern "C"

