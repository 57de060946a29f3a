// Fused Cyclix Hamiltonian stencil + Chebyshev three-term update (gfx950).
//
// out = alpha * ((H - shift) in) - beta * prev, with
//   H = -1/2 Lap + Veff + KB,
//   Lap = (D_rr + r^-1 D_r) + (tau^2 + r^-2) D_tt + D_zz + (-2 tau D_t) (x) D_z
// in (r, theta, z), radial Dirichlet, theta/z periodic with Bloch phases.
// This is h_nonlocal_vector_mult (src/h_nonlocal_vector_mult.m:21-43) over
// lapVec_cychel (Cychel/lapVec_cychel.m:19-53) folded into one step of
// chebyshev_filter (src/chebyshev_filter.m:29-37).
//
// Decomposition: one workgroup owns an r-tile x the whole theta ring of one
// column, and marches along z.  Each incoming z-plane is staged in LDS with
// its r halo and a theta halo whose wrapped copies already carry the Bloch
// phase; the z direction (and the z half of the helical cross term) is
// accumulated in a 13-deep register queue of output planes whose indices are
// compile-time constants (the plane loop is unrolled by 13).  An output plane
// is complete FDN planes after it entered the queue; it then takes the KB
// scatter (CSR by node), the recurrence, and is written in place over prev.
#include "cyx_internal.hpp"

namespace cyx {

// workgroup size cap: 512 threads (8 waves).  Types up to 8 bytes fit in
// 128 VGPRs, so two workgroups share a CU (4 waves per SIMD) and one's
// barrier waits overlap the other's work; complex128 needs ~220 VGPRs and
// runs one workgroup per CU.
template <class T>
constexpr int stencil_max_threads() {
  return 512;
}
template <class T>
constexpr int stencil_min_blocks() {
  return sizeof(T) >= 16 ? 1 : 2;
}

template <class T>
struct StepParams {
  using R = typename traits<T>::real;
  const T* in;
  const T* prev;
  T* out;
  int64_t ld;
  int nr, nt, nz;
  int tr, zb;
  int ntiles_r, nblocks_z;
  int xcd_remap;
  int nontemporal;
  R alpha, beta, shift;
  const R* rcoef;
  const R* veff;
  R w2[FDN + 1], w1[FDN + 1];  // FD weights (src/initialization.m:403-419)
  R s_r2, s_r1, s_z2, s_x;      // -1/2/dr^2, -1/2/dr, -1/2/dz^2, (-2 tau/dt)(-1/2/dz)
  R xin, dr;
  T ph_tl, ph_tr, ph_zl, ph_zr;
  int trh;  // LDS row stride (elements): tr + halo + bank padding
  const int* s_ptr;
  const R* s_chi;
  const int* s_il;
  const T* il_phase;
  const int* il_proj;
  int nimgl;
  const T* kbC;
  int nproj;
};

// KB entries per node held in registers (more are read from memory)
constexpr int KB_SLOTS = 5;

// non-temporal (streaming) element access: the X_prev stream and the result
// are touched once, so they should not displace halo planes and Veff in L2
__device__ __forceinline__ double ntload(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ntload(const float* p) { return __builtin_nontemporal_load(p); }
template <class R>
__device__ __forceinline__ cplx<R> ntload(const cplx<R>* p) {
  const R* q = reinterpret_cast<const R*>(p);
  return cplx<R>{__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1)};
}
__device__ __forceinline__ void ntstore(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void ntstore(float* p, float v) { __builtin_nontemporal_store(v, p); }
template <class R>
__device__ __forceinline__ void ntstore(cplx<R>* p, cplx<R> v) {
  R* q = reinterpret_cast<R*>(p);
  __builtin_nontemporal_store(v.x, q);
  __builtin_nontemporal_store(v.y, q + 1);
}

template <class T, bool HELICAL, bool HAS_PREV, bool HAS_KB>
__global__ void __launch_bounds__(stencil_max_threads<T>(), stencil_min_blocks<T>())
    stencil_step_kernel(const StepParams<T> P) {
  using R = typename traits<T>::real;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);

  const int TRH = P.trh;
  const int plane_sz = TRH * (P.nt + 2 * FDN);
  T* Dsh = lds + 2 * plane_sz;  // per-column KB table D[il]

  int bid = blockIdx.x;
  if (P.xcd_remap) {
    const int per = gridDim.x >> 3;
    bid = (bid & 7) * per + (bid >> 3);
  }
  const int tile = bid % P.ntiles_r;
  const int zblk = (bid / P.ntiles_r) % P.nblocks_z;
  const int s = bid / (P.ntiles_r * P.nblocks_z);

  const int tid = threadIdx.x;
  const int i = tid % P.tr;
  const int j = tid / P.tr;
  const int r0 = tile * P.tr;
  const int r = r0 + i;
  const bool rvalid = r < P.nr;
  const int z0 = zblk * P.zb;
  const int z1 = min(z0 + P.zb, P.nz);
  const int nrt = P.nr * P.nt;
  const int64_t colbase = (int64_t)s * P.ld;
  const T* __restrict__ src = P.in + colbase;
  const int jr = j * P.nr + r;  // in-plane node offset

  // per-radius scalars: (1/r) radial first-derivative factor, theta scale,
  // diagonal; radial taps are s_r2 sum w2_p (x+ + x-) + rfac sum w1_p (x+ - x-)
  R rfac = 0, cth = 0, cdg = 0;
  if (rvalid) {
    rfac = P.s_r1 / (P.xin + R(r) * P.dr);
    cth = P.rcoef[r * RCOEF + 2 * FDN];
    cdg = P.rcoef[r * RCOEF + 2 * FDN + 1] - P.shift;
  }

  if constexpr (HAS_KB) {
    // D[il] = e^{-i k.d_img} C[s][proj(il)] for this column
    for (int t = tid; t < P.nimgl; t += blockDim.x)
      Dsh[t] = cmul(P.il_phase[t], P.kbC[(int64_t)s * P.nproj + P.il_proj[t]]);
  }

  T q[QN];
#pragma unroll
  for (int u = 0; u < QN; ++u) q[u] = zero<T>();

  const int kbeg = z0 - FDN;
  const int nplanes = (z1 - z0) + 2 * FDN;

  // plane loader: virtual plane kk -> registers (with z Bloch phase)
  auto load_plane = [&](int kk, T& xa, T& xb, R& ve) {
    int kp = kk;
    int wrap = 0;
    if (kp < 0) {
      kp += P.nz;
      wrap = -1;
    } else if (kp >= P.nz) {
      kp -= P.nz;
      wrap = 1;
    }
    const T* row = src + (int64_t)kp * nrt + (int64_t)j * P.nr;
    const int ra = r0 - FDN + i;
    xa = (ra >= 0 && ra < P.nr) ? row[ra] : zero<T>();
    xb = zero<T>();
    if (i < 2 * FDN) {
      const int rb = r0 - FDN + P.tr + i;
      if (rb < P.nr) xb = row[rb];
    }
    if constexpr (traits<T>::is_complex) {
      if (wrap != 0) {
        const T ph = wrap < 0 ? P.ph_zl : P.ph_zr;
        xa = cmul(ph, xa);
        xb = cmul(ph, xb);
      }
    }
    ve = (rvalid && kk >= z0 && kk < z1) ? P.veff[(int64_t)kp * nrt + jr] : R(0);
  };

  // KB pipeline.  At the start of step kk: (hold, cur, nxt) are the CSR row
  // pointers of planes (kk, kk+1, kk+2) and echi/eil hold up to KB_SLOTS
  // entries of plane kk.  The step consumes plane kk's entries, issues the
  // entry loads of kk+1 and the row-pointer loads of kk+3.
  int sp_hold0 = 0, sp_hold1 = 0;
  int sp_cur0 = 0, sp_cur1 = 0;
  int sp_nxt0 = 0, sp_nxt1 = 0;
  R echi[KB_SLOTS];
  int eil[KB_SLOTS];
#pragma unroll
  for (int e = 0; e < KB_SLOTS; ++e) {
    echi[e] = R(0);
    eil[e] = 0;
  }
  T pv_cur = zero<T>(), pv_nxt = zero<T>();

  T xa, xb;
  R ve;
  load_plane(kbeg, xa, xb, ve);
  int buf = 0;
  if constexpr (HAS_KB) __syncthreads();  // Dsh ready

  for (int base = 0; base < nplanes; base += QN) {
#pragma unroll
    for (int u = 0; u < QN; ++u) {
      const int t = base + u;
      if (t < nplanes) {
        const int kk = kbeg + t;
        const bool inner = kk >= z0 && kk < z1;
        // ---- stage plane kk in LDS (theta halo carries the Bloch phase)
        T* L = lds + buf * plane_sz;
        {
          const int row = (j + FDN) * TRH;
          L[row + i] = xa;
          if (i < 2 * FDN) L[row + P.tr + i] = xb;
          if (j < FDN) {
            const int rw = (j + FDN + P.nt) * TRH;
            L[rw + i] = cmul(P.ph_tr, xa);
            if (i < 2 * FDN) L[rw + P.tr + i] = cmul(P.ph_tr, xb);
          }
          if (j >= P.nt - FDN) {
            const int rw = (j + FDN - P.nt) * TRH;
            L[rw + i] = cmul(P.ph_tl, xa);
            if (i < 2 * FDN) L[rw + P.tr + i] = cmul(P.ph_tl, xb);
          }
        }
        const R vcur = ve;
        // ---- KB of plane kk into its queue slot; then issue the next loads
        if constexpr (HAS_KB) {
          if (inner && rvalid) {
            T kb = zero<T>();
            const int cnt = sp_hold1 - sp_hold0;
#pragma unroll
            for (int e = 0; e < KB_SLOTS; ++e)
              if (e < cnt) kb = rfma(echi[e], Dsh[eil[e]], kb);
            for (int e = sp_hold0 + KB_SLOTS; e < sp_hold1; ++e) kb = rfma(P.s_chi[e], Dsh[P.s_il[e]], kb);
            q[u % QN] = add(q[u % QN], kb);
          }
          // entries of plane kk+1 (row pointers loaded one step ago)
          sp_hold0 = sp_cur0;
          sp_hold1 = sp_cur1;
          {
            const int cnt = sp_hold1 - sp_hold0;
#pragma unroll
            for (int e = 0; e < KB_SLOTS; ++e)
              if (e < cnt) {
                echi[e] = P.s_chi[sp_hold0 + e];
                eil[e] = P.s_il[sp_hold0 + e];
              }
          }
          // row pointers: (cur, nxt) become planes (kk+2, kk+3)
          sp_cur0 = sp_nxt0;
          sp_cur1 = sp_nxt1;
          sp_nxt0 = sp_nxt1 = 0;
          if (rvalid && kk + 3 >= z0 && kk + 3 < z1) {
            const int64_t n3 = (int64_t)(kk + 3) * nrt + jr;
            sp_nxt0 = P.s_ptr[n3];
            sp_nxt1 = P.s_ptr[n3 + 1];
          }
        }
        __syncthreads();
        // ---- prefetch: next plane, and X_prev of the plane emitted next step
        if (t + 1 < nplanes) load_plane(kk + 1, xa, xb, ve);
        const int ko = kk - FDN;
        const bool emit = (ko >= z0) && (ko < z1);
        if constexpr (HAS_PREV) {
          pv_cur = pv_nxt;
          if (ko + 1 >= z0 && ko + 1 < z1 && rvalid) {
            const T* pp = P.prev + colbase + (int64_t)(ko + 1) * nrt + jr;
            pv_nxt = P.nontemporal ? ntload(pp) : *pp;
          }
        }
        // ---- compute plane kk
        const int c = (j + FDN) * TRH + (i + FDN);
        const T x0 = L[c];
        T tt = zero<T>();
        T st = zero<T>();
#pragma unroll
        for (int p = 1; p <= FDN; ++p) {
          const T xp = L[c + p * TRH];
          const T xm = L[c - p * TRH];
          st = rfma(P.w2[p], add(xp, xm), st);
          if constexpr (HELICAL) tt = rfma(P.w1[p], sub(xp, xm), tt);
        }
        if (inner) {
          T sr = zero<T>();
          T dr = zero<T>();
#pragma unroll
          for (int p = 1; p <= FDN; ++p) {
            const T xp = L[c + p];
            const T xm = L[c - p];
            sr = rfma(P.w2[p], add(xp, xm), sr);
            dr = rfma(P.w1[p], sub(xp, xm), dr);
          }
          T self = rfma(P.s_r2, sr, rmul(cdg + vcur, x0));
          self = rfma(rfac, dr, self);
          self = rfma(cth, st, self);
          q[u % QN] = add(q[u % QN], self);
        }
        {
          const T xz = rmul(P.s_z2, x0);
          T tz = zero<T>();
          if constexpr (HELICAL) tz = rmul(P.s_x, tt);
#pragma unroll
          for (int d = 1; d <= FDN; ++d) {
            // plane kk feeds output planes kk-d and kk+d
            T& qa = q[(u - d + QN) % QN];
            T& qb = q[(u + d) % QN];
            qa = rfma(P.w2[d], xz, qa);
            qb = rfma(P.w2[d], xz, qb);
            if constexpr (HELICAL) {
              qa = rfma(P.w1[d], tz, qa);
              qb = rfma(-P.w1[d], tz, qb);
            }
          }
        }
        // ---- output plane kk-FDN is complete
        {
          T& qo = q[(u + FDN + 1) % QN];
          if (emit && rvalid) {
            T y = rmul(P.alpha, qo);
            if constexpr (HAS_PREV) y = rfma(-P.beta, pv_cur, y);
            T* po = P.out + colbase + (int64_t)ko * nrt + jr;
            if (P.nontemporal)
              ntstore(po, y);
            else
              *po = y;
          }
          qo = zero<T>();
        }
        buf ^= 1;
      }
    }
  }
}

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

// r tile: largest tile with tr * nt <= maxthr and tr >= 2*FDN, balanced over nr
static int choose_tr(int nr, int nt, int maxthr) {
  int maxtr = maxthr / nt;
  if (maxtr > nr) maxtr = nr;
  if (maxtr < 2 * FDN) maxtr = 2 * FDN;
  int ntiles = (nr + maxtr - 1) / maxtr;
  int tr = (nr + ntiles - 1) / ntiles;
  if (tr < 2 * FDN) tr = 2 * FDN;
  return tr;
}

template <class T, bool HEL, bool PREV, bool KB>
static void launch_variant(const StepParams<T>& P, int nblk, int nthr, size_t shm, hipStream_t st) {
  hipLaunchKernelGGL((stencil_step_kernel<T, HEL, PREV, KB>), dim3(nblk), dim3(nthr), shm, st, P);
}

template <class T>
static int step_impl(cyx_ham_s* h, const void* in_, const void* prev_, void* out_, int ncol, int64_t ld, double alpha,
                     double shift, double beta) {
  using R = typename traits<T>::real;
  const cyx_grid_desc& g = h->g;
  const OpCoef<R>& C = coef_of<R>(h);
  const KbTables<T>& K = kb_of<T>(h);
  hipStream_t st = h->ctx->stream;
  const int64_t N = h->N();
  const bool has_kb = K.nproj > 0;
  if (has_kb) {
    CYX_CHECK(ensure(h->kbC, (size_t)ncol * K.nproj * sizeof(T)) == CYX_OK, "kb scratch allocation failed");
    PhaseTimer tm(h->ctx, PH_KB_GATHER);
    hipLaunchKernelGGL(kb_gather_kernel<T>, dim3(K.nproj, ncol), dim3(256), 0, st, (const T*)in_, ld, K.g_ptr,
                       K.g_node, K.g_coef, K.gamma, K.nproj, (T*)h->kbC.p);
    CYX_HIP(hipGetLastError());
  }
  StepParams<T> P{};
  P.in = (const T*)in_;
  P.prev = (const T*)prev_;
  P.out = (T*)out_;
  P.ld = ld;
  P.nr = g.nr;
  P.nt = g.nt;
  P.nz = g.nz;
  P.tr = choose_tr(g.nr, g.nt, stencil_max_threads<T>());
  if (h->tune_tr > 0) P.tr = std::max(2 * FDN, std::min(h->tune_tr, stencil_max_threads<T>() / g.nt));
  P.ntiles_r = ceil_div(g.nr, P.tr);
  int zb = g.nz;
  while ((long)P.ntiles_r * ceil_div(g.nz, zb) * ncol < 512 && zb > 4) zb = (zb + 1) / 2;
  if (h->tune_zb > 0) zb = std::min(h->tune_zb, g.nz);
  P.zb = zb;
  P.nontemporal = (h->tune_flags & 1) ? 1 : 0;
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
  P.xin = (R)g.xin;
  P.dr = (R)g.dr;
  P.ph_tl = phase_of<T>(h->ph[0][0], h->ph[0][1]);
  P.ph_tr = phase_of<T>(h->ph[1][0], h->ph[1][1]);
  P.ph_zl = phase_of<T>(h->ph[2][0], h->ph[2][1]);
  P.ph_zr = phase_of<T>(h->ph[3][0], h->ph[3][1]);
  P.s_ptr = K.s_ptr;
  P.s_chi = K.s_chi;
  P.s_il = K.s_il;
  P.il_phase = K.il_phase;
  P.il_proj = K.il_proj;
  P.nimgl = K.nimgl;
  P.kbC = (const T*)h->kbC.p;
  P.nproj = K.nproj;
  // LDS row stride: halo, then padding so that the next row starts on the
  // next bank group (row pitch minus tile width a multiple of 256 B)
  P.trh = P.tr + std::max(2 * FDN, (int)(256 / sizeof(T)));
  const long nblk = (long)P.ntiles_r * P.nblocks_z * ncol;
  CYX_CHECK(nblk < (1L << 31), "grid too large");
  P.xcd_remap = (nblk % 8 == 0 && !(h->tune_flags & 2)) ? 1 : 0;
  const int nthr = P.tr * g.nt;
  const size_t shm = (2 * (size_t)P.trh * (g.nt + 2 * FDN) + (size_t)K.nimgl) * sizeof(T);
  const bool hel = g.cell_typ == 4 || g.cell_typ == 5;
  const bool prev = prev_ != nullptr;
  (void)N;
#define CYX_LAUNCH(HEL, PRV, KBV) launch_variant<T, HEL, PRV, KBV>(P, (int)nblk, nthr, shm, st)
  PhaseTimer tm(h->ctx, prev ? PH_STENCIL : PH_HAPPLY);
  if (hel) {
    if (prev) {
      if (has_kb) CYX_LAUNCH(true, true, true); else CYX_LAUNCH(true, true, false);
    } else {
      if (has_kb) CYX_LAUNCH(true, false, true); else CYX_LAUNCH(true, false, false);
    }
  } else {
    if (prev) {
      if (has_kb) CYX_LAUNCH(false, true, true); else CYX_LAUNCH(false, true, false);
    } else {
      if (has_kb) CYX_LAUNCH(false, false, true); else CYX_LAUNCH(false, false, false);
    }
  }
#undef CYX_LAUNCH
  CYX_HIP(hipGetLastError());
  return CYX_OK;
}

int stencil_step(cyx_ham_s* h, int dtype, const void* in, const void* prev, void* out, int ncol, int64_t ld,
                 double alpha, double shift, double beta) {
  switch (dtype) {
    case CYX_F64:
      return step_impl<double>(h, in, prev, out, ncol, ld, alpha, shift, beta);
    case CYX_F32:
      return step_impl<float>(h, in, prev, out, ncol, ld, alpha, shift, beta);
    case CYX_C128:
      return step_impl<cplx<double>>(h, in, prev, out, ncol, ld, alpha, shift, beta);
    case CYX_C64:
      return step_impl<cplx<float>>(h, in, prev, out, ncol, ld, alpha, shift, beta);
  }
  set_error("unknown dtype %d", dtype);
  return CYX_ERR_ARG;
}

}  // namespace cyx
