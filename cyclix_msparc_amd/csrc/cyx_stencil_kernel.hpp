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
//
// This header holds the launch parameters (always) and, when
// CYX_STENCIL_KERNELS is defined, the kernels and their launchers; each
// cyx_stencil_<type>.hip translation unit instantiates the launchers of one
// storage type, so the variants compile in parallel.
#pragma once
#include <type_traits>

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
// z planes staged per barrier: complex128 (one workgroup per CU) takes two,
// so that the LDS reads of one plane overlap the FMAs of the other and the
// barriers halve; smaller types keep one and run two workgroups per CU
template <class T>
constexpr int stencil_pp() {
  return sizeof(T) >= 16 ? 2 : 1;
}
// radial tile of the compile-time-shaped variants (the heuristic tile of the
// 120- and 80-radius grids at nt = 24 and of the 40-radius grid at nt = 16)
constexpr int STENCIL_TR = 20;
// LDS row stride (elements) for radial tile tr: the tile, its halo, and
// padding so that the next row starts on the next bank group (row pitch
// minus tile width a multiple of 256 bytes)
template <class T>
constexpr int stencil_trh(int tr) {
  return tr + (2 * FDN > (int)(256 / sizeof(T)) ? 2 * FDN : (int)(256 / sizeof(T)));
}
// Wide theta rings: a workgroup holds tr x nt <= 512 threads, so rings of
// more than 42 nodes take radial tiles narrower than the 2 FDN halo; a row's
// thread i then loads positions i + h tr (h = 1..STENCIL_NH_WIDE) of the
// tile's extended row.  Radial tiles down to STENCIL_TR_MIN (nt <= 128).
constexpr int STENCIL_TR_MIN = 4;
static_assert(MAX_NT * STENCIL_TR_MIN <= 512, "theta ring limit and narrowest tile disagree");
constexpr int STENCIL_NH_WIDE = (2 * FDN + STENCIL_TR_MIN - 1) / STENCIL_TR_MIN;
constexpr int STENCIL_WIDE = -1;  // the NT value of the narrow-tile variants
// the compile-time theta count of a launch: 24 or 16 with the tile
// STENCIL_TR, STENCIL_WIDE for radial tiles narrower than the halo, else 0
// (shape read at run time)
inline int stencil_nt_variant(int nt, int tr) {
  if (tr < 2 * FDN) return STENCIL_WIDE;
  return (tr == STENCIL_TR && (nt == 24 || nt == 16)) ? nt : 0;
}

// T: arithmetic type of the step; TS: storage type of in/prev/out (T, or
// the FP32 form of T for an FP64 step over an FP32 block)
template <class T, class TS = T>
struct StepParams {
  using R = typename traits<T>::real;
  const TS* in;
  const TS* prev;
  TS* out;
  int64_t ld;
  int nr, nt, nz;
  int tr, zb;
  int ntiles_r, nblocks_z;
  int xcd_remap;
  int nontemporal;
  int tile_major, ncol;
  R alpha, beta, shift;
  const R* rcoef;
  const R* veff;
  R w2[FDN + 1], w1[FDN + 1];  // FD weights (src/initialization.m:403-419)
  R s_r2, s_r1, s_z2, s_x;      // -1/2/dr^2, -1/2/dr, -1/2/dz^2, (-2 tau/dt)(-1/2/dz)
  R wz2[FDN + 1], wx1[FDN + 1];  // s_z2 w2 and s_x w1: the z-direction and cross-term weights
  R xin, dr;
  T ph_tl, ph_tr, ph_zl, ph_zr;
  int trh;  // LDS row stride (elements): tr + halo + bank padding
  const uint32_t* s_pk;
  const R* s_chi;
  const int* s_il;
  const T* il_phase;
  const int* il_proj;
  int nimgl;
  const T* kbC;
  int nproj;
};

// Several channels (Bloch phases, grids, column counts) in one launch: entry
// c owns blocks [E[c].blk0, E[c+1].blk0).  All entries share nt and the r
// tile (one workgroup shape) and the template variant.
template <class T>
struct BatchEntry {
  StepParams<T> P;
  int blk0, nblk;
};

// KB entries per node held in registers (more are read from memory)
constexpr int KB_SLOTS = 5;

// load/store conversion between the storage type and the arithmetic type
template <class T, class TS>
__device__ __forceinline__ T to_arith(TS v) {
  if constexpr (std::is_same_v<T, TS>)
    return v;
  else
    return widen(v);
}
template <class TS, class T>
__device__ __forceinline__ TS to_store(T v) {
  if constexpr (std::is_same_v<T, TS>)
    return v;
  else
    return narrow<TS>(v);
}

// Launch one recurrence step (kernel variant chosen by hel / prev / kb);
// defined in the translation unit of the storage type TS.  An FP64 step
// over an FP32 block (T wider than TS) has no prev operand.
template <class T, class TS>
void launch_step(const StepParams<T, TS>& P, int nblk, int nthr, size_t shm, bool hel, bool prev, bool kb,
                 hipStream_t st);
// Launch one batched step over several channels (T storage = arithmetic).
// nt_variant: stencil_nt_variant of the channels' common (nt, tile).
template <class T>
void launch_batch(const BatchEntry<T>* E, int nch, int nblk, int nthr, size_t shm, bool hel, bool prev, bool kb,
                  int nt_variant, hipStream_t st);

}  // namespace cyx

#ifdef CYX_STENCIL_KERNELS
namespace cyx {

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


// The body of one workgroup: block `bid0` of the `nblk` blocks of one
// channel's launch (the batched kernel runs several channels in one grid).
// NT > 0: the theta count is NT and the radial tile STENCIL_TR, both known at
// compile time, so every LDS tap is an immediate offset from one address;
// NT = 0 reads them from P; NT = STENCIL_WIDE also, with radial tiles below
// 2 FDN (several right-halo loads per thread).
template <class T, class TS, int NT, bool HELICAL, bool HAS_PREV, bool HAS_KB>
__device__ __forceinline__ void stencil_body(const StepParams<T, TS>& P, int bid0, int nblk, T* lds) {
  using R = typename traits<T>::real;
  constexpr int PP = stencil_pp<T>();
  // right-halo registers per thread: positions i + h tr_ of the extended row
  // [r0 - FDN, r0 + tr_ + FDN), h = 1..NH
  constexpr int NH = NT == STENCIL_WIDE ? STENCIL_NH_WIDE : 1;
  // radial taps from per-thread merged coefficients (FP32 arithmetic: 12
  // registers; FP64 keeps the shared-weight form to save 24 registers)
  constexpr bool MERGED = sizeof(R) == 4;
  const int nt_ = NT > 0 ? NT : P.nt;
  const int tr_ = NT > 0 ? STENCIL_TR : P.tr;
  const int TRH = NT > 0 ? stencil_trh<T>(STENCIL_TR) : P.trh;
  const int plane_sz = TRH * (nt_ + 2 * FDN);
  T* Dsh = lds + 2 * PP * plane_sz;  // per-column KB table D[il]

  int bid = bid0;
  if (P.xcd_remap) {
    const int per = nblk >> 3;
    bid = (bid & 7) * per + (bid >> 3);
  }
  // block order: radial tile fastest, or (tile-major) the column fastest so
  // that the blocks resident on one XCD share a tile's Veff, coefficients and
  // projector tables in its L2
  int tile, zblk, s;
  if (P.tile_major) {
    const int per_tile = P.nblocks_z * P.ncol;
    tile = bid / per_tile;
    const int rem = bid - tile * per_tile;
    zblk = rem % P.nblocks_z;
    s = rem / P.nblocks_z;
  } else {
    tile = bid % P.ntiles_r;
    zblk = (bid / P.ntiles_r) % P.nblocks_z;
    s = bid / (P.ntiles_r * P.nblocks_z);
  }

  const int tid = threadIdx.x;
  const int i = tid % tr_;
  const int j = tid / tr_;
  const int r0 = tile * tr_;
  const int r = r0 + i;
  const bool rvalid = r < P.nr;
  const int z0 = zblk * P.zb;
  const int z1 = min(z0 + P.zb, P.nz);
  const int nrt = P.nr * nt_;
  const int64_t colbase = (int64_t)s * P.ld;
  const TS* __restrict__ src = P.in + colbase;
  const int jr = j * P.nr + r;  // in-plane node offset

  // per-radius scalars: (1/r) radial first-derivative factor, theta scale,
  // diagonal; radial taps are s_r2 sum w2_p (x+ + x-) + rfac sum w1_p (x+ - x-)
  R rfac = 0, cth = 0, cdg = 0;
  if (rvalid) {
    rfac = P.s_r1 / (P.xin + R(r) * P.dr);
    cth = P.rcoef[r * RCOEF + 2 * FDN];
    cdg = P.rcoef[r * RCOEF + 2 * FDN + 1] - P.shift;
  }
  // merged radial coefficients: s_r2 w2_p +- rfac w1_p for the taps r +- p
  R cpl[FDN], cmi[FDN];
#pragma unroll
  for (int p = 1; p <= FDN; ++p) {
    cpl[p - 1] = rfma(rfac, P.w1[p], P.s_r2 * P.w2[p]);
    cmi[p - 1] = rfma(-rfac, P.w1[p], P.s_r2 * P.w2[p]);
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
  auto load_plane = [&](int kk, T& xa, T (&xb)[NH], R& ve) {
    int kp = kk;
    int wrap = 0;
    if (kp < 0) {
      kp += P.nz;
      wrap = -1;
    } else if (kp >= P.nz) {
      kp -= P.nz;
      wrap = 1;
    }
    const TS* row = src + (int64_t)kp * nrt + (int64_t)j * P.nr;
    const int ra = r0 - FDN + i;
    xa = (ra >= 0 && ra < P.nr) ? to_arith<T>(row[ra]) : zero<T>();
#pragma unroll
    for (int hh = 0; hh < NH; ++hh) {
      xb[hh] = zero<T>();
      const int pos = i + (hh + 1) * tr_;
      if (pos < tr_ + 2 * FDN) {
        const int rb = r0 - FDN + pos;
        if ((NH == 1 || rb >= 0) && rb < P.nr) xb[hh] = to_arith<T>(row[rb]);
      }
    }
    if constexpr (traits<T>::is_complex) {
      if (wrap != 0) {
        const T ph = wrap < 0 ? P.ph_zl : P.ph_zr;
        xa = cmul(ph, xa);
#pragma unroll
        for (int hh = 0; hh < NH; ++hh) xb[hh] = cmul(ph, xb[hh]);
      }
    }
    ve = (rvalid && kk >= z0 && kk < z1) ? P.veff[(int64_t)kp * nrt + jr] : R(0);
  };

  if constexpr (PP == 1) {
  // KB pipeline.  At the start of step kk: (hold, cur, nxt) are the CSR row
  // pointers of planes (kk, kk+1, kk+2) and echi/eil hold up to KB_SLOTS
  // entries of plane kk.  The step consumes plane kk's entries, issues the
  // entry loads of kk+1 and the row-pointer loads of kk+3.
  uint32_t sp_hold = 0, sp_cur = 0, sp_nxt = 0;  // packed (first << 6) | count
  R echi[KB_SLOTS];
  int eil[KB_SLOTS];
#pragma unroll
  for (int e = 0; e < KB_SLOTS; ++e) {
    echi[e] = R(0);
    eil[e] = 0;
  }
  T pv_cur = zero<T>(), pv_nxt = zero<T>();

  T xa, xb[NH];
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
#pragma unroll
          for (int hh = 0; hh < NH; ++hh)
            if (i + hh * tr_ < 2 * FDN) L[row + tr_ + hh * tr_ + i] = xb[hh];
          if (j < FDN) {
            const int rw = (j + FDN + nt_) * TRH;
            L[rw + i] = cmul(P.ph_tr, xa);
#pragma unroll
            for (int hh = 0; hh < NH; ++hh)
              if (i + hh * tr_ < 2 * FDN) L[rw + tr_ + hh * tr_ + i] = cmul(P.ph_tr, xb[hh]);
          }
          if (j >= nt_ - FDN) {
            const int rw = (j + FDN - nt_) * TRH;
            L[rw + i] = cmul(P.ph_tl, xa);
#pragma unroll
            for (int hh = 0; hh < NH; ++hh)
              if (i + hh * tr_ < 2 * FDN) L[rw + tr_ + hh * tr_ + i] = cmul(P.ph_tl, xb[hh]);
          }
        }
        const R vcur = ve;
        // ---- KB of plane kk into its queue slot; then issue the next loads
        if constexpr (HAS_KB) {
          if (inner && rvalid) {
            T kb = zero<T>();
            const int h0 = (int)(sp_hold >> 6), cnt = (int)(sp_hold & 63u);
#pragma unroll
            for (int e = 0; e < KB_SLOTS; ++e)
              if (e < cnt) kb = rfma(echi[e], Dsh[eil[e]], kb);
            for (int e = h0 + KB_SLOTS; e < h0 + cnt; ++e) kb = rfma(P.s_chi[e], Dsh[P.s_il[e]], kb);
            q[u % QN] = add(q[u % QN], kb);
          }
          // entries of plane kk+1 (row pointers loaded one step ago)
          sp_hold = sp_cur;
          {
            const int h0 = (int)(sp_hold >> 6), cnt = (int)(sp_hold & 63u);
#pragma unroll
            for (int e = 0; e < KB_SLOTS; ++e)
              if (e < cnt) {
                echi[e] = P.s_chi[h0 + e];
                eil[e] = P.s_il[h0 + e];
              }
          }
          // node words: (cur, nxt) become planes (kk+2, kk+3)
          sp_cur = sp_nxt;
          sp_nxt = 0;
          if (rvalid && kk + 3 >= z0 && kk + 3 < z1) sp_nxt = P.s_pk[(int64_t)(kk + 3) * nrt + jr];
        }
        __syncthreads();
        // ---- prefetch: next plane, and X_prev of the plane emitted next step
        if (t + 1 < nplanes) load_plane(kk + 1, xa, xb, ve);
        const int ko = kk - FDN;
        const bool emit = (ko >= z0) && (ko < z1);
        if constexpr (HAS_PREV) {
          pv_cur = pv_nxt;
          if (ko + 1 >= z0 && ko + 1 < z1 && rvalid) {
            const TS* pp = P.prev + colbase + (int64_t)(ko + 1) * nrt + jr;
            pv_nxt = to_arith<T>(P.nontemporal ? ntload(pp) : *pp);
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
          T self = rfma(cth, st, rmul(cdg + vcur, x0));
          if constexpr (MERGED) {
#pragma unroll
            for (int p = 1; p <= FDN; ++p) {
              self = rfma(cpl[p - 1], L[c + p], self);
              self = rfma(cmi[p - 1], L[c - p], self);
            }
          } else {
            T sr = zero<T>();
            T dr = zero<T>();
#pragma unroll
            for (int p = 1; p <= FDN; ++p) {
              const T xp = L[c + p];
              const T xm = L[c - p];
              sr = rfma(P.w2[p], add(xp, xm), sr);
              dr = rfma(P.w1[p], sub(xp, xm), dr);
            }
            self = rfma(P.s_r2, sr, self);
            self = rfma(rfac, dr, self);
          }
          q[u % QN] = add(q[u % QN], self);
        }
#pragma unroll
        for (int d = 1; d <= FDN; ++d) {
          // plane kk feeds output planes kk-d and kk+d (z weights carry
          // -1/2/dz^2 and the helical cross factor)
          T& qa = q[(u - d + QN) % QN];
          T& qb = q[(u + d) % QN];
          qa = rfma(P.wz2[d], x0, qa);
          qb = rfma(P.wz2[d], x0, qb);
          if constexpr (HELICAL) {
            qa = rfma(P.wx1[d], tt, qa);
            qb = rfma(-P.wx1[d], tt, qb);
          }
        }
        // ---- output plane kk-FDN is complete
        {
          T& qo = q[(u + FDN + 1) % QN];
          if (emit && rvalid) {
            T y = rmul(P.alpha, qo);
            if constexpr (HAS_PREV) y = rfma(-P.beta, pv_cur, y);
            TS* po = P.out + colbase + (int64_t)ko * nrt + jr;
            if (P.nontemporal)
              ntstore(po, to_store<TS>(y));
            else
              *po = to_store<TS>(y);
          }
          qo = zero<T>();
        }
        buf ^= 1;
      }
    }
  }
  } else {

  // ---- two planes per barrier (see stencil_pp)
  // stage a plane: interior row, its r halo, and the theta halo copies with
  // the Bloch phase
  auto stage = [&](T* L, const T& xa, const T (&xb)[NH]) {
    const int row = (j + FDN) * TRH;
    L[row + i] = xa;
#pragma unroll
    for (int hh = 0; hh < NH; ++hh)
      if (i + hh * tr_ < 2 * FDN) L[row + tr_ + hh * tr_ + i] = xb[hh];
    if (j < FDN) {
      const int rw = (j + FDN + nt_) * TRH;
      L[rw + i] = cmul(P.ph_tr, xa);
#pragma unroll
      for (int hh = 0; hh < NH; ++hh)
        if (i + hh * tr_ < 2 * FDN) L[rw + tr_ + hh * tr_ + i] = cmul(P.ph_tr, xb[hh]);
    }
    if (j >= nt_ - FDN) {
      const int rw = (j + FDN - nt_) * TRH;
      L[rw + i] = cmul(P.ph_tl, xa);
#pragma unroll
      for (int hh = 0; hh < NH; ++hh)
        if (i + hh * tr_ < 2 * FDN) L[rw + tr_ + hh * tr_ + i] = cmul(P.ph_tl, xb[hh]);
    }
  };
  // KB pipeline, per plane slot p of a pair (planes t0+p): hold[p] is the
  // node word of plane t0+p with its entries in echi2[p]/eil2[p]; nw[p] the
  // word of plane t0+p+2.  Planes z0-6 .. z0-3 carry no entries, so the
  // zero start is exact.
  uint32_t holdA = 0u, holdB = 0u, nwA = 0u, nwB = 0u;
  R echiA[KB_SLOTS], echiB[KB_SLOTS];
  int eilA[KB_SLOTS], eilB[KB_SLOTS];
#pragma unroll
  for (int e = 0; e < KB_SLOTS; ++e) {
    echiA[e] = echiB[e] = R(0);
    eilA[e] = eilB[e] = 0;
  }
  // one plane's KB step: the scatter sum of plane kk (returned), then the
  // entry loads of plane kk+2 and the node word of plane kk+4
  auto kb_step = [&](int kk, uint32_t& hold, uint32_t& nw, R(&ech)[KB_SLOTS], int(&eix)[KB_SLOTS]) -> T {
    T kb = zero<T>();
    if (kk >= z0 && kk < z1 && rvalid) {
      const int h0 = (int)(hold >> 6), cnt = (int)(hold & 63u);
#pragma unroll
      for (int e = 0; e < KB_SLOTS; ++e)
        if (e < cnt) kb = rfma(ech[e], Dsh[eix[e]], kb);
      for (int e = h0 + KB_SLOTS; e < h0 + cnt; ++e) kb = rfma(P.s_chi[e], Dsh[P.s_il[e]], kb);
    }
    hold = nw;
    {
      const int h0 = (int)(hold >> 6), cnt = (int)(hold & 63u);
#pragma unroll
      for (int e = 0; e < KB_SLOTS; ++e)
        if (e < cnt) {
          ech[e] = P.s_chi[h0 + e];
          eix[e] = P.s_il[h0 + e];
        }
    }
    nw = 0u;
    if (rvalid && kk + 4 >= z0 && kk + 4 < z1) nw = P.s_pk[(int64_t)(kk + 4) * nrt + jr];
    return kb;
  };
  // X_prev of the two planes emitted next (z0 and z0+1 first)
  T pv2[2] = {zero<T>(), zero<T>()};
  if constexpr (HAS_PREV) {
    if (rvalid) {
      const TS* pp = P.prev + colbase + (int64_t)z0 * nrt + jr;
      pv2[0] = to_arith<T>(P.nontemporal ? ntload(pp) : *pp);
      if (z0 + 1 < z1) pv2[1] = to_arith<T>(P.nontemporal ? ntload(pp + nrt) : pp[nrt]);
    }
  }
  T xa2[2], xb2[2][NH];
  R ve2[2];
  load_plane(kbeg, xa2[0], xb2[0], ve2[0]);
  xa2[1] = zero<T>();
#pragma unroll
  for (int hh = 0; hh < NH; ++hh) xb2[1][hh] = zero<T>();
  ve2[1] = R(0);
  if (nplanes > 1) load_plane(kbeg + 1, xa2[1], xb2[1], ve2[1]);
  int buf = 0;
  if constexpr (HAS_KB) __syncthreads();  // Dsh ready

  for (int base = 0; base < nplanes; base += 2 * QN) {
#pragma unroll
    for (int v = 0; v < QN; ++v) {
      const int t0 = base + 2 * v;
      if (t0 < nplanes) {
        const int kk0 = kbeg + t0;
        T* Lp[2] = {lds + (2 * buf) * plane_sz, lds + (2 * buf + 1) * plane_sz};
        R vc[2];
        // the two planes of the pair, with the plane index a compile-time
        // constant (the queue slots must be static register indices)
        auto stage_plane = [&](auto pc) {
          constexpr int p = decltype(pc)::value;
          if (t0 + p < nplanes) {
            stage(Lp[p], xa2[p], xb2[p]);
            vc[p] = ve2[p];
            if constexpr (HAS_KB) {
              const T kb = p == 0 ? kb_step(kk0, holdA, nwA, echiA, eilA) : kb_step(kk0 + 1, holdB, nwB, echiB, eilB);
              q[(2 * v + p) % QN] = add(q[(2 * v + p) % QN], kb);
            }
          }
        };
        stage_plane(std::integral_constant<int, 0>{});
        stage_plane(std::integral_constant<int, 1>{});
        __syncthreads();
        // ---- prefetch the next pair
        if (t0 + 2 < nplanes) load_plane(kk0 + 2, xa2[0], xb2[0], ve2[0]);
        if (t0 + 3 < nplanes) load_plane(kk0 + 3, xa2[1], xb2[1], ve2[1]);
        auto compute_plane = [&](auto pc) {
          constexpr int p = decltype(pc)::value;
          if (t0 + p < nplanes) {
            const int u = 2 * v + p;
            const int kk = kk0 + p;
            const bool inner = kk >= z0 && kk < z1;
            const T* L = Lp[p];
            const int c = (j + FDN) * TRH + (i + FDN);
            const T x0 = L[c];
            T tt = zero<T>();
            T st = zero<T>();
#pragma unroll
            for (int d = 1; d <= FDN; ++d) {
              const T xp = L[c + d * TRH];
              const T xm = L[c - d * TRH];
              st = rfma(P.w2[d], add(xp, xm), st);
              if constexpr (HELICAL) tt = rfma(P.w1[d], sub(xp, xm), tt);
            }
            if (inner) {
              T sr = zero<T>();
              T dr = zero<T>();
#pragma unroll
              for (int d = 1; d <= FDN; ++d) {
                const T xp = L[c + d];
                const T xm = L[c - d];
                sr = rfma(P.w2[d], add(xp, xm), sr);
                dr = rfma(P.w1[d], sub(xp, xm), dr);
              }
              T self = rfma(P.s_r2, sr, rmul(cdg + vc[p], x0));
              self = rfma(rfac, dr, self);
              self = rfma(cth, st, self);
              q[u % QN] = add(q[u % QN], self);
            }
#pragma unroll
            for (int d = 1; d <= FDN; ++d) {
              T& qa = q[(u - d + QN) % QN];
              T& qb = q[(u + d) % QN];
              qa = rfma(P.wz2[d], x0, qa);
              qb = rfma(P.wz2[d], x0, qb);
              if constexpr (HELICAL) {
                qa = rfma(P.wx1[d], tt, qa);
                qb = rfma(-P.wx1[d], tt, qb);
              }
            }
            // output plane kk-FDN is complete; emit it before the next plane
            // of the pair reuses its queue slot
            const int ko = kk - FDN;
            T& qo = q[(u + QN - FDN) % QN];
            if (ko >= z0 && ko < z1 && rvalid) {
              T y = rmul(P.alpha, qo);
              if constexpr (HAS_PREV) y = rfma(-P.beta, pv2[p], y);
              TS* po = P.out + colbase + (int64_t)ko * nrt + jr;
              if (P.nontemporal)
                ntstore(po, to_store<TS>(y));
              else
                *po = to_store<TS>(y);
              if constexpr (HAS_PREV) {
                if (ko + 2 < z1) {
                  const TS* pp = P.prev + colbase + (int64_t)(ko + 2) * nrt + jr;
                  pv2[p] = to_arith<T>(P.nontemporal ? ntload(pp) : *pp);
                }
              }
            }
            qo = zero<T>();
          }
        };
        compute_plane(std::integral_constant<int, 0>{});
        compute_plane(std::integral_constant<int, 1>{});
        buf ^= 1;
      }
    }
  }
  }
}

// waves per SIMD: 2 for complex128, 4 (two 8-wave workgroups per CU) otherwise
template <class T>
constexpr int stencil_waves_per_eu() {
  return 2 * stencil_min_blocks<T>();
}

template <class T, class TS, int NT, bool HELICAL, bool HAS_PREV, bool HAS_KB>
__global__ void __launch_bounds__(stencil_max_threads<T>())
    __attribute__((amdgpu_waves_per_eu(stencil_waves_per_eu<T>()))) stencil_step_kernel(const StepParams<T, TS> P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  stencil_body<T, TS, NT, HELICAL, HAS_PREV, HAS_KB>(P, blockIdx.x, gridDim.x, reinterpret_cast<T*>(smem_raw));
}

template <class T, int NT, bool HELICAL, bool HAS_PREV, bool HAS_KB>
__global__ void __launch_bounds__(stencil_max_threads<T>())
    __attribute__((amdgpu_waves_per_eu(stencil_waves_per_eu<T>())))
    stencil_batch_kernel(const BatchEntry<T>* __restrict__ E, int nch) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int c = 0;
  while (c + 1 < nch && E[c + 1].blk0 <= (int)blockIdx.x) ++c;
  stencil_body<T, T, NT, HELICAL, HAS_PREV, HAS_KB>(E[c].P, blockIdx.x - E[c].blk0, E[c].nblk,
                                                 reinterpret_cast<T*>(smem_raw));
}

// dynamic LDS above 64 KB (two-plane staging of complex128) must be allowed
// per kernel; gfx950 has 160 KB per CU
inline void allow_lds(const void* fn, size_t shm) {
  if (shm > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
}

template <class T, class TS, int NT, bool HEL, bool PREV, bool KB>
void launch_variant(const StepParams<T, TS>& P, int nblk, int nthr, size_t shm, hipStream_t st) {
  allow_lds((const void*)&stencil_step_kernel<T, TS, NT, HEL, PREV, KB>, shm);
  hipLaunchKernelGGL((stencil_step_kernel<T, TS, NT, HEL, PREV, KB>), dim3(nblk), dim3(nthr), shm, st, P);
}

// the helical / prev choices of one compile-time shape NT (widening steps
// have no prev operand)
template <class T, class TS, int NT, bool KB>
void launch_shape(const StepParams<T, TS>& P, int nblk, int nthr, size_t shm, bool hel, bool prev, hipStream_t st) {
  if constexpr (std::is_same_v<T, TS>) {
    if (hel) {
      if (prev) launch_variant<T, TS, NT, true, true, KB>(P, nblk, nthr, shm, st);
      else launch_variant<T, TS, NT, true, false, KB>(P, nblk, nthr, shm, st);
    } else {
      if (prev) launch_variant<T, TS, NT, false, true, KB>(P, nblk, nthr, shm, st);
      else launch_variant<T, TS, NT, false, false, KB>(P, nblk, nthr, shm, st);
    }
  } else {
    (void)prev;
    if (hel) launch_variant<T, TS, NT, true, false, KB>(P, nblk, nthr, shm, st);
    else launch_variant<T, TS, NT, false, false, KB>(P, nblk, nthr, shm, st);
  }
}

template <class T, class TS>
void launch_step(const StepParams<T, TS>& P, int nblk, int nthr, size_t shm, bool hel, bool prev, bool kb,
                 hipStream_t st) {
  // the KB term inside the plane loop exists in the run-time-shaped variant
  // only (the host keeps it out of narrow-tile launches)
  if (kb) return launch_shape<T, TS, 0, true>(P, nblk, nthr, shm, hel, prev, st);
  switch (stencil_nt_variant(P.nt, P.tr)) {
    case 24:
      return launch_shape<T, TS, 24, false>(P, nblk, nthr, shm, hel, prev, st);
    case 16:
      return launch_shape<T, TS, 16, false>(P, nblk, nthr, shm, hel, prev, st);
    case STENCIL_WIDE:
      return launch_shape<T, TS, STENCIL_WIDE, false>(P, nblk, nthr, shm, hel, prev, st);
    default:
      return launch_shape<T, TS, 0, false>(P, nblk, nthr, shm, hel, prev, st);
  }
}

template <class T, int NT, bool HEL, bool PREV, bool KB>
void launch_batch_variant(const BatchEntry<T>* E, int nch, int nblk, int nthr, size_t shm, hipStream_t st) {
  allow_lds((const void*)&stencil_batch_kernel<T, NT, HEL, PREV, KB>, shm);
  hipLaunchKernelGGL((stencil_batch_kernel<T, NT, HEL, PREV, KB>), dim3(nblk), dim3(nthr), shm, st, E, nch);
}

template <class T, int NT>
void launch_batch_shape(const BatchEntry<T>* E, int nch, int nblk, int nthr, size_t shm, bool hel, bool prev, bool kb,
                        hipStream_t st) {
#define CYX_BL(HEL, PRV, KBV) launch_batch_variant<T, NT, HEL, PRV, KBV>(E, nch, nblk, nthr, shm, st)
  if (hel) {
    if (prev) {
      if (kb) CYX_BL(true, true, true); else CYX_BL(true, true, false);
    } else {
      if (kb) CYX_BL(true, false, true); else CYX_BL(true, false, false);
    }
  } else {
    if (prev) {
      if (kb) CYX_BL(false, true, true); else CYX_BL(false, true, false);
    } else {
      if (kb) CYX_BL(false, false, true); else CYX_BL(false, false, false);
    }
  }
#undef CYX_BL
}

template <class T>
void launch_batch(const BatchEntry<T>* E, int nch, int nblk, int nthr, size_t shm, bool hel, bool prev, bool kb,
                  int nt_variant, hipStream_t st) {
  if (nt_variant == 24) return launch_batch_shape<T, 24>(E, nch, nblk, nthr, shm, hel, prev, kb, st);
  return launch_batch_shape<T, 0>(E, nch, nblk, nthr, shm, hel, prev, kb, st);
}

}  // namespace cyx
#endif  // CYX_STENCIL_KERNELS
