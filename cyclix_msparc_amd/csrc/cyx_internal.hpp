// Internal (non-ABI) declarations shared by the library's translation units.
#pragma once
#include <vector>

#include "../../include/cyclix_msparc.h"
#include "cyx_common.hpp"

struct rocblas_handle_s;

namespace cyx {

constexpr int FDN = 6;            // finite-difference half-width the kernels are built for
constexpr int QN = 2 * FDN + 1;   // z queue depth of the marching stencil
constexpr int RCOEF = 2 * FDN + 2;  // per-radius coefficients: c+[6], c-[6], c_theta, c_diag
constexpr int MAX_NT = 128;        // theta ring limit of the marching stencil (512 threads, radial tiles >= 4)

// order-2 FDN central weights, src/initialization.m:403-419: w2 second and
// w1 first derivative, index 0 the centre
inline void host_fd_weights(double* w2, double* w1) {
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
}

// Per-channel operator coefficients in one precision (R = double or float).
template <class R>
struct OpCoef {
  R* rcoef = nullptr;   // [nr][RCOEF]   (-1/2 folded in)
  R* veff = nullptr;    // [N]
  R w2[FDN + 1];        // FD weights of the second derivative
  R w1[FDN + 1];        // FD weights of the first derivative
  R s_r2, s_r1, s_z2;   // -1/2/dr^2, -1/2/dr (times 1/r), -1/2/dz^2
  R s_x;                // (-2 tau/dt)(-1/2/dz): helical cross term
};

// KB tables in one arithmetic type T (complex phases for Bloch channels).
template <class T>
struct KbTables {
  using R = typename traits<T>::real;
  int nproj = 0;
  // gather, sorted by projector: C[s][p] = gamma_p sum_e g[e] * Y[node[e]]
  int* g_ptr = nullptr;   // [nproj+1]
  int* g_node = nullptr;  // [nent]
  T* g_coef = nullptr;    // [nent]  chi * W * e^{+i k.d}
  R* gamma = nullptr;     // [nproj]
  // scatter, CSR by node: HY[n] += sum_e chi[e] * D[il[e]], where for one
  // column D[il] = e^{-i k.d_img} * C[proj(il)] over (image, projector) pairs
  uint32_t* s_pk = nullptr;  // [N] (first entry << 6) | entry count (count <= 63)
  R* s_chi = nullptr;     // [nent]
  int* s_il = nullptr;    // [nent] (image, projector) pair index
  int nimgl = 0;
  T* il_phase = nullptr;  // [nimgl] e^{-i k.d_img}
  int* il_proj = nullptr; // [nimgl]
  int* kb_nodes = nullptr;  // [nkbn] the nodes with at least one entry, ascending
  int nkbn = 0;
  T* s_cph = nullptr;       // [nent] chi * e^{-i k.d_img}, for the standalone scatter
  int* s_pj = nullptr;      // [nent] projector of the entry
  // two-stage gather over node-range chunks shared by all projectors (each
  // projector's entries ascend by node): pair q = (projector, chunk) covers
  // that projector's entries [pr_lo[q], pr_hi[q]) in the chunk; block x takes
  // pair blk_pair[x] (-1: none), so that the pairs of one chunk run on one
  // XCD back to back; pairs are numbered projector-major in chunk order, so
  // projector p owns pairs [pp_ptr[p], pp_ptr[p+1])
  int npair = 0, nblk_pair = 0;
  int* pr_lo = nullptr;     // [npair]
  int* pr_hi = nullptr;     // [npair]
  int* blk_pair = nullptr;  // [nblk_pair], a multiple of 8
  int* pp_ptr = nullptr;    // [nproj+1]
};

struct DeviceBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace cyx

namespace cyx {
// Phases timed with HIP events on the context's stream when timing is on.
enum Phase {
  PH_STENCIL = 0,
  PH_HAPPLY,
  PH_KB_GATHER,
  PH_GRAM,
  PH_ROTATE,
  PH_WNORM,
  PH_DENSE_EIG,
  PH_AAR,
  PH_RR_APPLY,
  PH_STERN,
  PH_COUNT
};
}

struct cyx_ctx_s {
  int device = 0;
  int ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* blas = nullptr;  // rocblas_handle
  bool timing = false;
  struct Mark {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;            // recorded, not yet resolved
  double phase_ms[cyx::PH_COUNT] = {};
  int64_t phase_count[cyx::PH_COUNT] = {};
};

namespace cyx {
// RAII event pair around one device phase (no-op when timing is off)
struct PhaseTimer {
  cyx_ctx_s* c;
  int phase;
  hipEvent_t a = nullptr;
  PhaseTimer(cyx_ctx_s* ctx, int ph) : c(ctx), phase(ph) {
    if (c->timing && hipEventCreate(&a) == hipSuccess) (void)hipEventRecord(a, c->stream);
  }
  ~PhaseTimer() {
    if (c->timing && a) {
      hipEvent_t b = nullptr;
      if (hipEventCreate(&b) == hipSuccess) {
        (void)hipEventRecord(b, c->stream);
        c->marks.push_back({phase, a, b});
      }
    }
  }
};
}  // namespace cyx

namespace cyx {
// exact inverse of Lap_std (k = 0) by a theta/z FFT and one band solve per mode
struct Spectral;
int spectral_create(const cyx_grid_desc& g, Spectral** out);
void spectral_destroy(Spectral* S);
// z = Lap_std^-1 r for ncol columns (leading dimension ld) on stream st
int spectral_solve_real(Spectral* S, const double* r, double* z, int ncol, int64_t ld, hipStream_t st);
int spectral_solve_complex(Spectral* S, const cplx<double>* r, cplx<double>* z, int ncol, int64_t ld, hipStream_t st);
}  // namespace cyx

struct cyx_ham_s {
  cyx_ctx_s* ctx = nullptr;
  int device = 0;
  cyx_grid_desc g{};
  double kpt[3] = {0, 0, 0};
  bool bloch = false;     // any non-zero Bloch phase (requires complex dtypes)
  double* W = nullptr;    // [N] FP64 quadrature weights S.W (normalisation, KB gather)
  double* Wm = nullptr;   // [N] metric weights r dr dt dz (H is self-adjoint in them)
  cyx::OpCoef<double> c64;
  cyx::OpCoef<float> c32;
  // phases e^{-i nu L2}, e^{+i nu L2}, e^{-i eta L3}, e^{+i eta L3}
  double ph[4][2];
  cyx::KbTables<double> kb_d;
  cyx::KbTables<float> kb_f;
  cyx::KbTables<cyx::cplx<double>> kb_z;
  cyx::KbTables<cyx::cplx<float>> kb_c;
  int nproj = 0;
  // stencil launch shape (0: heuristic), see cyx_tuning
  int tune_tr = 0, tune_zb = 0, tune_flags = 0;
  // Rayleigh-Ritz rotations of FP32 blocks on the FP32 / FP64 matrix cores
  int64_t rot32_count = 0, rot64_count = 0;
  // lazily grown scratch
  cyx::DeviceBuf kbC;      // C coefficients (ncol x nproj x 16 B)
  cyx::DeviceBuf small;    // Gram partial sums
  cyx::DeviceBuf sub;      // M, Hs, eigenvalues, solver work/info, column scales
  cyx::DeviceBuf red;      // per-block partial sums of the energy and SCF-error reductions
  cyx::DeviceBuf stern;    // Sternheimer operator / BiCGSTAB workspace
  cyx::Spectral* spec = nullptr;  // exact Lap_std^-1 of the grid (Sternheimer preconditioner), built on first use
  int64_t N() const { return (int64_t)g.nr * g.nt * g.nz; }
};

struct cyx_poisson_s {
  cyx_ctx_s* ctx = nullptr;
  cyx_unet_s* net = nullptr;  // learned preconditioner (null: Jacobi only)
  cyx::Spectral* spec = nullptr;  // exact Lap_std^-1 (built by cyx_poisson_set_spectral)
  bool use_spectral = false;      // AAR preconditioned by spec instead of the Jacobi diagonal
  cyx_grid_desc g{};
  double* coef = nullptr;   // [nr][RCOEF] Lap_std coefficients (no -1/2)
  double* coefT = nullptr;  // [RCOEF][nr] the same, radius fastest, taps outside [0, nr) zeroed
  double* scratch = nullptr;
  size_t scratch_bytes = 0;
  cyx::DeviceBuf tbuf;      // theta-derivative buffer of cyx_poisson_apply
  int64_t N() const { return (int64_t)g.nr * g.nt * g.nz; }
};

namespace cyx {

int ensure(DeviceBuf& b, size_t bytes);

// Caller-provided KB scratch of a stencil step: C (ncol x nproj of the
// arithmetic type) and the gather partials (ncol x npair, FP64 accumulator
// type).  Null: the handle's lazily grown buffers.
struct KbScratch {
  void* C;
  void* part;
};
// bytes of KB scratch one step of `ncol` columns needs (C region rounded up
// to 256 bytes, then the partials)
size_t kb_scratch_bytes(cyx_ham_s* h, int dtype, int ncol);
// the no-allocation entry points' device scratch: a file-scope device array
// of NOALLOC_SCRATCH_BYTES, its address looked up once per device
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
int noalloc_scratch_init(int device);   // with `device` current
void* noalloc_scratch(int device);      // null before init

// stencil step: out = alpha*((H - shift) in) - beta*prev   (prev may be null;
// out may alias prev).  Dispatches on dtype.  Includes the KB term when the
// handle has projectors (gather launched internally on `in`).
int stencil_step(cyx_ham_s* h, int dtype, const void* in, const void* prev, void* out, int ncol, int64_t ld,
                 double alpha, double shift, double beta, const KbScratch* ks = nullptr);
// out = H in with FP64 arithmetic and the FP64 coefficients whatever the
// storage type of in/out (FP32 blocks are widened on load and rounded on
// store): the Rayleigh-Ritz and residual H-applies of the mixed path
int stencil_apply_fp64(cyx_ham_s* h, int dtype, const void* in, void* out, int ncol, int64_t ld);
// Y = alpha (V X + KB X): the handle's potential and projector terms without
// the kinetic stencil (X: ldx, Y: ldy, Y != X)
int potential_apply(cyx_ham_s* h, int dtype, const void* X, int64_t ldx, void* Y, int64_t ldy, int ncol,
                    double alpha, const KbScratch* ks = nullptr);

// subspace kernels (FP64 accumulation)
int gram_w(cyx_ham_s* h, int dtype, const void* A, const void* B, int ns, int64_t ld, void* M_out /*dev*/);
// M (na x nb, FP64 / complex FP64, column-major) = A^H diag(W) B over the N grid points;
// herm: the product is Hermitian (na == nb), only the upper tiles are computed and mirrored
// part: device scratch of gram_part_bytes bytes for the partial sums (null:
// the handle's lazily grown buffer; captured graphs pass their own)
int gram_rect(cyx_ham_s* h, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
              const double* W, void* M_out /*dev*/, bool herm = false, void* part = nullptr);
size_t gram_part_bytes(int dtype, int64_t N, int na, int nb, bool herm);
int rotate(cyx_ham_s* h, int dtype, const void* A, const void* Q /*dev fp64 ns x ns*/, void* out, int ns, int64_t ld);
// the same products of FP32 blocks on the FP32 matrix cores (cyx_subspace32.hip):
// the Gram with FP32 sums over 128-point chunks added in FP64; the rotation
// with Q rounded to FP32 and FP32 sums (for well-conditioned blocks only)
int gram_rect32(cyx_ham_s* h, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                const double* W, void* M_out /*dev*/, bool herm);
int rotate32(cyx_ham_s* h, int dtype, const void* A, const void* Q /*dev fp64 ns x ns*/, void* out, int ns, int64_t ld);
// out[i] = Re M_ii (device); ratio = min_i R_ii^2 / M_ii for the Cholesky
// factor R stored over M (device scalar)
int overlap_diag(cyx_ham_s* h, bool cx, const void* M, int ns, double* out);
int cholesky_ratio(cyx_ham_s* h, bool cx, const void* R, const double* mdiag, int ns, double* out);
int wnormalize(cyx_ham_s* h, int dtype, void* A, int ns, int64_t ld, double* scal /*dev ns*/);
int hermitize(cyx_ham_s* h, int dtype, void* M, int ns);
// res[s] = ||HA_s - lam_s A_s||_W / ||A_s||_W in FP64 (lam: device, ns)
int ritz_residuals(cyx_ham_s* h, int dtype, const void* A, const void* HA, const double* lam, int ns, int64_t ld,
                   double* res /*dev ns*/);

// learned preconditioner: u = U(D^{-1} r) (FP32), see cyx_unet.hip
int unet_precondition(cyx_unet_s* net, const double* coef, const double* r, float* u);
bool unet_matches(const cyx_unet_s* net, const cyx_grid_desc& g);
cyx_ctx_s* unet_context(const cyx_unet_s* net);

// Householder reduction of a Hermitian (cx) or symmetric FP64 n x n matrix to
// real tridiagonal form in one workgroup (LAPACK lower storage of the
// reflectors); hetrd_lds_bytes(n) <= 64 KiB is required
size_t hetrd_lds_bytes(int n, bool cx);
int hetrd(hipStream_t st, bool cx, void* A, int n, double* d, double* e, void* tau);
// The same reduction over up to 32 workgroups with the matrix in LDS (one
// grid barrier per column); 0 groups: n does not fit.  ws holds
// hetrd_dist_ws_bytes(n) bytes; *flag points at a device word that is
// nonzero after the launch if a barrier wait gave up.
int hetrd_dist_groups(int n, bool cx);
size_t hetrd_dist_ws_bytes(int n, bool cx);
int hetrd_dist(hipStream_t st, bool cx, void* A, int n, double* d, double* e, void* tau, void* ws,
               unsigned** flag);

// Poisson kernels
int poisson_run(cyx_poisson_s* ps, const double* f, double* x, int use_guess, double tol, int max_iter,
                double omega, double beta, int m, int p, cyx_aar_info* info);
int poisson_rhs(cyx_poisson_s* ps, const double* rho, const double* b, double* f);
int poisson_apply(cyx_poisson_s* ps, const double* x, double* y);

// SCF glue
int density_accumulate(cyx_ham_s* h, int dtype, const void* psi, int ns, const double* occ_dev, double c, double* rho);
int lda_pw(hipStream_t st, int64_t N, const double* rho, const double* phi, double* vxc, double* veff, double rhotol);
int density_response(cyx_ham_s* h, int dtype, const void* psi, const void* dpsi, int ns, const double* coef_dev,
                     double c, double* drho);
int lda_pw_fxc(hipStream_t st, int64_t N, const double* rho, const double* drho, const double* dphi, double* fxc,
               double* dveff, double rhotol);
int energy_terms(hipStream_t st, int64_t N, const double* rho, const double* vxc, const double* phi, const double* b,
                 const double* W, double rhotol, double* part_dev, int nblk, double out[3]);
int veff_from_device(cyx_ham_s* h, const double* veff);

size_t dtype_size(int dtype);

// Sternheimer plans (cyx_sternheimer.hip): enqueue `iters` BiCGSTAB
// iterations on st (the context's stream is switched to st and timing is off
// meanwhile, so that the launches can be captured into a caller's graph)
int stern_plan_enqueue(cyx_stern_plan_s* sp, hipStream_t st, int iters);
cyx_ctx_s* stern_plan_context(const cyx_stern_plan_s* sp);
// whether the plan preconditions with the exact Lap_std^-1 (its iterations are not captured)
bool stern_plan_spectral(const cyx_stern_plan_s* sp);

}  // namespace cyx
