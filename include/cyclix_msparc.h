/*
 * cyclix_msparc.h -- C-ABI of the MI355X (gfx950) CheFSI + AAR-Poisson library.
 *
 * Drop-in boundary for the hot path of Cyclix-M-SPARC's SCF step.  A MATLAB
 * host binds these symbols from a MEX gateway; every entry point names the
 * reference routine it replaces (paths relative to Cyclix_MS_200220/).
 *
 * Conventions
 *  - Grid functions are length-N vectors, N = nr*nt*nz, radial index fastest,
 *    then theta, then z (src/calculate_nloc_projector.m:145).
 *  - A block of states is column-major N x ncol with leading dimension ld >= N
 *    (MATLAB's psi(:,:,ks), src/scf.m:82).  Complex data is interleaved
 *    (re, im), the MATLAB R2018a interleaved-complex layout.
 *  - Pointers documented "device" are HIP device pointers owned by the caller;
 *    "host" pointers are ordinary host memory.  No torch types cross the ABI.
 *  - Every function returns CYX_OK (0) or a negative status; cyx_last_error()
 *    returns a message for the last failure on the calling thread.  No C++
 *    exception crosses the ABI.
 *  - All work is enqueued on the context's HIP stream; functions that return
 *    host-visible results (eigenvalues, AAR info) synchronise that stream.
 *  - A context (and every handle made from it) is single-caller: calls that
 *    overlap on one handle race.
 */
#ifndef CYCLIX_MSPARC_H
#define CYCLIX_MSPARC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CYX_OK 0
#define CYX_ERR_ARG (-1)      /* invalid argument or shape */
#define CYX_ERR_HIP (-2)      /* HIP runtime error */
#define CYX_ERR_NOCONV (-3)   /* iterative solver hit max_iter (result still returned) */
#define CYX_ERR_NOMEM (-4)    /* device allocation failed */
#define CYX_ERR_SOLVER (-5)   /* dense solver failure (e.g. overlap not positive definite) */
#define CYX_ERR_STATE (-6)    /* handle used in a wrong state */

/* arithmetic types of a block of states */
#define CYX_F64 0  /* real double      (Gamma channel, FP64)            */
#define CYX_C128 1 /* complex double   (Bloch channel, FP64)            */
#define CYX_F32 2  /* real float       (mixed precision, Gamma)         */
#define CYX_C64 3  /* complex float    (mixed precision, Bloch channel) */

/* One cyclic/helical sector.  Fields are the reference's S.Nx, S.Ny, S.Nz
 * (node counts incl. the Dirichlet node, src/initialization.m:341-344),
 * S.FDn, S.cell_typ (3 cyclic, 4/5 helical), S.xin, S.dx, S.dy, S.dz, S.twist.
 * Radial direction Dirichlet, theta and z periodic with Bloch phases. */
typedef struct cyx_grid_desc {
  int32_t nr, nt, nz;
  int32_t fdn;
  int32_t cell_typ;
  int32_t reserved;
  double xin, dr, dt, dz;
  double twist;
} cyx_grid_desc;

/* Kleinman-Bylander projectors: S.Atom(J).rcImage(img).{rc_pos,Chi_mat,
 * coordinates} and S.Atom(J).gamma_Jl flattened over atoms and images
 * (src/calculate_nloc_projector.m:185-277).  All host pointers. */
typedef struct cyx_projector_desc {
  int32_t n_atom;               /* S.n_atm                                            */
  int32_t n_image;              /* rc-images summed over atoms                        */
  const int32_t* angnum;        /* [n_atom] projectors per atom (S.Atom(J).angnum)     */
  const double* gamma;          /* [sum angnum] gamma_Jl, atoms concatenated           */
  const int32_t* image_atom;    /* [n_image] 0-based owning atom                       */
  const int32_t* image_nnodes;  /* [n_image] rc_n_nodes                                */
  const double* image_disp;     /* [3*n_image] S.Atoms(J,:) - rcImage(img).coordinates */
  const int32_t* rc_pos;        /* [sum nnodes] 0-based node index, images concatenated */
  const double* chi;            /* per image, Chi_mat column-major nnodes x angnum      */
} cyx_projector_desc;

typedef struct cyx_ctx_s* cyx_ctx_t;
typedef struct cyx_ham_s* cyx_ham_t;
typedef struct cyx_poisson_s* cyx_poisson_t;
typedef struct cyx_mixer_s* cyx_mixer_t;
typedef struct cyx_unet_s* cyx_unet_t;
typedef struct cyx_filter_plan_s* cyx_filter_plan_t;
typedef struct cyx_stern_plan_s* cyx_stern_plan_t;

/* ------------------------------------------------------------ context */
const char* cyx_last_error(void);
const char* cyx_version(void);
/* Creates a context on HIP device `device` with its own (default-priority)
 * stream.  Replaces nothing in the reference (MATLAB owns no device). */
int cyx_ctx_create(int device, cyx_ctx_t* out);
int cyx_ctx_destroy(cyx_ctx_t ctx);
/* Use a caller-owned HIP stream (hipStream_t) instead of the context's. */
int cyx_ctx_set_stream(cyx_ctx_t ctx, void* hip_stream);
int cyx_ctx_synchronize(cyx_ctx_t ctx);
/* Per-phase device timers (HIP events recorded on the context's stream
 * around every launch of a phase).  Phases: "stencil" (fused H-apply +
 * three-term Chebyshev step), "h_apply" (stencil launches without the
 * three-term update: first filter step, Lanczos), "kb_gather" (KB gathers
 * and scatters), "gram", "rotate", "wnorm", "dense_eig", "aar" (one whole
 * AAR solve), "rr_apply" (the FP64-arithmetic H psi of Rayleigh-Ritz and of
 * the residual check), "sternheimer" (one replay of a Sternheimer plan's
 * captured iterations).  The reference prints only tic/toc wall times
 * (src/scf.m:284-304, src/poissonSolve.m:49). */
int cyx_ctx_set_timing(cyx_ctx_t ctx, int enable);
int cyx_ctx_reset_timing(cyx_ctx_t ctx);
int cyx_ctx_read_timing(cyx_ctx_t ctx, const char* phase, double* total_ms, int64_t* count);

/* ------------------------------------------------------------ Hamiltonian */
/* Device-resident operator of one Bloch channel: the 1-D factors of
 * blochLaplacian_1d (src/blochLaplacian_1d.m:1-123 over
 * Cychel/lapIndicesValues_1d_cychel.m:9-261), S.W, S.Veff and the KB tables.
 * kpt = S.kptgrid(kpt,:) = [0, nu, eta].  veff, W: host, length N. */
int cyx_ham_create(cyx_ctx_t ctx, const cyx_grid_desc* grid, const double kpt[3], const double* veff,
                   const double* W, const cyx_projector_desc* proj, cyx_ham_t* out);
/* Replace Veff (S.Veff(:,spin)); `on_device` selects whether veff is a device pointer. */
int cyx_ham_set_veff(cyx_ham_t ham, const double* veff, int on_device);
int cyx_ham_destroy(cyx_ham_t ham);
/* Device pointer of the FP64 weights W held by the handle (length N). */
const double* cyx_ham_weights(cyx_ham_t ham);
/* How many Rayleigh-Ritz rotations of FP32 blocks ran on the FP32 matrix
 * cores (well-conditioned overlap) and how many on the FP64 ones, since the
 * handle was made.  No counterpart in the reference. */
int cyx_ham_rotation_counts(cyx_ham_t ham, int64_t* fp32, int64_t* fp64);

/* Launch shape of the fused stencil kernel (the autotuner's knobs; no
 * counterpart in the reference).  tr: radial tile (0 = heuristic; clamped to
 * the kernel's thread limit), zb: z planes per workgroup (0 = heuristic),
 * flags: bit 0 non-temporal loads of X_prev and stores of the result, bit 1
 * disable the XCD-aware block remap, bit 2 tile-major block order (blocks of
 * one radial tile adjacent, so an XCD's resident blocks share that tile's
 * tables in L2).  Results do not depend on these. */
typedef struct cyx_tuning {
  int32_t tr, zb, flags, reserved;
} cyx_tuning;
int cyx_ham_set_tuning(cyx_ham_t ham, const cyx_tuning* t);

/* HX = H X.  Replaces h_nonlocal_vector_mult (src/h_nonlocal_vector_mult.m:1-43).
 * X, HX device, column-major N x ncol, leading dimension ld. */
int cyx_h_apply(cyx_ham_t ham, int dtype, const void* X, void* HX, int ncol, int64_t ld);

/* Y = p_m(H) X, Zhou-Saad scaled Chebyshev filter of degree m on [a, b] with
 * lower bound a0.  Replaces chebyshev_filter (src/chebyshev_filter.m:1-42).
 * X and Y device N x ncol (leading dimension ld); X is overwritten (used as
 * the recurrence's second buffer). */
int cyx_chebyshev_filter(cyx_ham_t ham, int dtype, void* X, void* Y, int ncol, int64_t ld, int m, double a,
                         double b, double a0);

/* cyx_chebyshev_filter without any device allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).  The same kernels run;
 * the scratch they need (the KB projection coefficients and gather partials,
 * ncol * (nproj * sizeof(element) + npair * 16) bytes) lives in a
 * library-owned device array of 16 MiB per GPU, held until the process ends.
 * A call whose scratch would exceed it fails with CYX_ERR_ARG before anything
 * is enqueued.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (from any context or thread) race; the array's contents
 * are fully rewritten before they are read, so no call depends on the
 * previous one.  A context must have been created on the device first. */
int cyx_chebyshev_filter_noalloc(cyx_ham_t ham, int dtype, void* X, void* Y, int ncol, int64_t ld, int m, double a,
                                 double b, double a0);

/* Bytes of device workspace cyx_eigsolver_pass needs for `ns` states. */
int64_t cyx_eigsolver_workspace_bytes(cyx_ham_t ham, int dtype, int ns);

/* One CheFSI pass of one channel: the body of eigSolver's channel loop
 * (src/eigSolver.m:98-124): filter of degree npl on [lambda_cutoff, bup] with
 * lower bound a0, then W-metric Rayleigh-Ritz (overlap Cholesky and projected
 * Hermitian eigenproblem in FP64 on the device), rotation and W-normalisation.
 * psi: device N x ns (ld = N), in/out.  work: device, at least
 * cyx_eigsolver_workspace_bytes.  eigval: host, ns doubles, ascending. */
int cyx_eigsolver_pass(cyx_ham_t ham, int dtype, void* psi, void* work, int ns, int npl, double lambda_cutoff,
                       double bup, double a0, double* eigval);

/* Rayleigh-Ritz only (src/eigSolver.m:99-124) on an already filtered psi. */
int cyx_rayleigh_ritz(cyx_ham_t ham, int dtype, void* psi, void* work, int ns, double* eigval);

/* FP64 residual check of Ritz pairs: resid[s] = ||H psi_s - eigval[s] psi_s||_W
 * / ||psi_s||_W with H applied in FP64 arithmetic and FP64 coefficients for
 * every storage type (an FP32 block is widened exactly), W = S.W.  It checks
 * the eigenpairs that eigSolver returns (src/eigSolver.m:110-124); the
 * reference computes no residual.  psi, work: device N x ns (ld = N);
 * eigval, resid: host, ns doubles. */
int cyx_ritz_residuals(cyx_ham_t ham, int dtype, const void* psi, void* work, int ns, const double* eigval,
                       double* resid);

/* M = A^H diag(W) B over the grid (W the handle's quadrature weights):
 * the PsiW * X products of the reference, e.g. the nonlocal force integrals
 * transpose(conj(S.psi) .* S.W) * [Chi_full, DChi_x, DChi_y, DChi_z]
 * (src/atomicForce.m:304-312).  A: device na columns (lda), B: device nb
 * columns (ldb), both of dtype; M: device, na x nb column-major, FP64 or
 * complex FP64 (accumulated in FP64).  Synchronises. */
int cyx_weighted_gram(cyx_ham_t ham, int dtype, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                      void* M);

/* Extreme Ritz values of H by Lanczos from start vector v0 (host, N doubles,
 * real part used); stands in for eigs(Hfun,N,1,'lr'/'sr') at
 * src/eigSolver.m:47-63 (tolerance TOL_LANCZOS, at most maxit steps). */
int cyx_lanczos_bounds(cyx_ham_t ham, const double* v0, int maxit, double tol, double* lambda_min,
                       double* lambda_max);

/* ------------------------------------------------------------ batched filter */
/* Chebyshev filters of many channels at once: the per-channel loop of
 * src/eigSolver.m:26-98 (chebyshev_filter on every (nu, eta) channel) as
 * one launch per recurrence step over all channels, the m steps of a pass
 * captured once into a HIP graph and replayed.  Channel c filters the
 * ncol[c] columns of A[c] (device, leading dimension ld[c]) with B[c] as
 * work; after a run the result is in B[c] when *result_in_b, else in A[c]
 * (A[c] is overwritten either way).  All channels share dtype, nt, the
 * helical/cyclic form and whether they have projectors; grids, Bloch phases
 * and column counts may differ.  autotune != 0 times one step per candidate
 * radial tile and block order on these buffers (using bounds a, b, a0) and
 * keeps the fastest (cyx_filter_plan_tuning reports it).  Buffers and
 * channels are fixed for the plan's life; bounds change per run. */
int cyx_filter_plan_create(cyx_ctx_t ctx, int nch, const cyx_ham_t* hams, int dtype, void* const* A, void* const* B,
                           const int* ncol, const int64_t* ld, int m, int autotune, const double* a, const double* b,
                           const double* a0, cyx_filter_plan_t* out);
/* one pass: bounds per channel (host arrays of nch), src/chebyshev_filter.m:24-42 */
int cyx_filter_plan_run(cyx_filter_plan_t plan, const double* a, const double* b, const double* a0,
                        int* result_in_b);
int cyx_filter_plan_tuning(cyx_filter_plan_t plan, cyx_tuning* t);
/* Puts `iters` iterations of a Sternheimer plan (declared further down,
 * same context) into the plan's graph beside the filter steps, as an
 * independent branch: every cyx_filter_plan_run then filters all channels
 * and advances all right-hand sides in one graph launch (DFPT solves
 * batched with the CheFSI channels).  sp = NULL detaches.  The Sternheimer
 * plan must outlive the attachment; poll it with
 * cyx_sternheimer_plan_status.  No counterpart in the reference, which runs
 * the two loops one after the other on the host. */
int cyx_filter_plan_attach_sternheimer(cyx_filter_plan_t plan, cyx_stern_plan_t sp, int iters);
int cyx_filter_plan_destroy(cyx_filter_plan_t plan);

/* ------------------------------------------------------------ SCF glue (FP64) */
/* rho += 2 * weight * sum_s occ[s] |psi_s|^2 over the ns device states psi
 * (ld = N).  Replaces one k-point term of electronDensity
 * (src/electronDensity.m:13-17).  occ: host, ns doubles; rho: device. */
int cyx_density_accumulate(cyx_ham_t ham, int dtype, const void* psi, int ns, const double* occ, double weight,
                           double* rho);
/* vxc = LDA Perdew-Wang potential of rho with rho clamped at rhotol
 * (src/exchangeCorrelationPotential.m:82-106); if veff is non-null also
 * veff = phi + vxc (src/scf.m:39).  All device, length N. */
int cyx_lda_pw(cyx_ctx_t ctx, int64_t N, const double* rho, const double* phi, double* vxc, double* veff,
               double rhotol);
/* DFPT density response at q = 0 (electronDensity_pert,
 * src/phonon.m:1232-1240): drho += 2 weight sum_s (dg[s] |psi_s|^2
 * + 2 occ[s] Re(conj(psi_s) dpsi_s)) over ns states; psi and dpsi are
 * (ns, N) device blocks of one dtype, occ and dg host arrays of ns (dg may
 * be null: 0, an insulator); drho an FP64 device vector of N. */
int cyx_density_response(cyx_ham_t ham, int dtype, const void* psi, const void* dpsi, int ns, const double* occ,
                         const double* dg, double weight, double* drho);
/* DFPT exchange-correlation kernel S.Dvxc_drho of the LDA Perdew-Wang
 * potential (src/phonon.m:215,303; the reference has it only for LDA_PZ via
 * libxc, src/exchangeCorrelationPotential.m:125): the closed-form derivative
 * of cyx_lda_pw's potential, 0 where rho < rhotol.  fxc (optional) receives
 * the kernel; dveff (optional) = dphi + fxc .* drho (src/phonon.m:303-306).
 * All device, length N. */
int cyx_lda_pw_fxc(cyx_ctx_t ctx, int64_t N, const double* rho, const double* drho, const double* dphi, double* fxc,
                   double* dveff, double rhotol);
/* out = {Exc, Exc_dc, Eelec_dc}: the grid sums of evaluateTotalEnergy
 * (src/evaluateTotalEnergy.m:27-52,72) with the handle's weights S.W. */
int cyx_energy_terms(cyx_ham_t ham, const double* rho, const double* vxc, const double* phi, const double* b,
                     double rhotol, double out[3]);
/* out = {Escc, ||vout - vin||^2, ||vout||^2} for potential mixing: the
 * self-consistency correction sum((Veff - Veff_temp) .* rho .* W)
 * (src/scf.m:160-162) and the two squared norms of the SCF error
 * norm(Veff - Veff_temp) / norm(Veff) (src/scf.m:182-184). */
int cyx_scf_terms(cyx_ham_t ham, const double* vout, const double* vin, const double* rho, double out[3]);

/* Periodic-Pulay mixing of a length-N FP64 device vector
 * (src/mixing.m:20-116 with MixingPrecond = 0, PulayRestartFreq = 0).
 * create: history m in [1, 8] (S.MixingHistory), beta (S.MixingParameter),
 * omega (S.MixingParameterSimple), p (S.PulayFrequency); x0 is the initial
 * input iterate, stored as x_{k-1} with f_{k-1} = 0 and X = F = 0
 * (src/scf.m:42-56).  step: x_out = mixing(g = output iterate, x = input
 * iterate, iter), iter 1-based as count_SCF; x_out may alias g but not x. */
int cyx_mixer_create(cyx_ctx_t ctx, int64_t N, int history, double beta, double omega, int pulay_frequency,
                     const double* x0, cyx_mixer_t* out);
int cyx_mixer_destroy(cyx_mixer_t mixer);
int cyx_mixer_step(cyx_mixer_t mixer, const double* g, const double* x, double* x_out, int iter);

/* ------------------------------------------------------------ Sternheimer (DFPT) */
/* Y_j = (H - lam[j]) X_j + psi_v (psi_v^H (W .* X_j)) for the ncol columns of
 * X: the operator H_times_x of src/phonon.m:2139-2157 (one shift per column,
 * H the handle's channel operator including the KB term, W the handle's
 * weights).  X, Y: device blocks (ncol x ld), psi_v: device block of nv
 * occupied states (nv x ldv; nv = 0 drops the projector term); lam: host,
 * ncol doubles. */
int cyx_sternheimer_apply(cyx_ham_t ham, int dtype, const void* X, void* Y, int ncol, int64_t ld, const double* lam,
                          const void* psi_v, int64_t ldv, int nv);

typedef struct cyx_stern_info {
  int32_t max_iterations; /* largest per-column iteration count */
  int32_t n_converged;    /* columns with ||b - A x|| <= tol ||b|| (recursive residual) */
  double max_relres;      /* largest per-column relative residual */
} cyx_stern_info;

/* X_j = bicgstab(A_j, B_j, tol, max_iter, M, X_j) for every column at once,
 * A_j the operator above: replaces the per-state bicgstab loop at
 * src/phonon.m:1516-1521.  M, where the reference passes the ILU factors
 * of S.Lap_std (src/initialization.m:752): precond = 2 the exact Lap_std^-1
 * (theta/z FFT and one band solve per mode, as cyx_poisson_set_spectral;
 * FP64 blocks; iterations run without graph capture), 1 the Jacobi diagonal
 * of -1/2 Lap (per radius), 0 none.  use_guess = 0 starts from X = 0.
 * Per-column iteration counts and relative residuals go to col_iters and
 * col_relres (host, may be NULL).  Returns CYX_ERR_NOCONV (with X and info
 * filled) when a column misses tol within max_iter or breaks down. */
int cyx_sternheimer_solve(cyx_ham_t ham, int dtype, const void* B, void* X, int ncol, int64_t ld, const double* lam,
                          const void* psi_v, int64_t ldv, int nv, int use_guess, double tol, int max_iter,
                          int precond, cyx_stern_info* info, int32_t* col_iters, double* col_relres);

/* The same solve as a resident plan whose iterations are captured in a HIP
 * graph: create (workspace, x0, r = b - A x0), then replay `iters`
 * iterations per call until cyx_sternheimer_plan_status reports no active
 * column.  Converged columns freeze on the device (their x and r stay), so a
 * replay never needs the host.  B, X, lam's columns and psi_v must stay
 * allocated and in place for the plan's life (the graph holds their
 * addresses).  Replaces the same per-state bicgstab loop
 * (src/phonon.m:1516-1521); cyx_sternheimer_solve runs on it. */
int cyx_sternheimer_plan_create(cyx_ham_t ham, int dtype, const void* B, void* X, int ncol, int64_t ld,
                                const double* lam, const void* psi_v, int64_t ldv, int nv, int use_guess, double tol,
                                int precond, cyx_stern_plan_t* out);
int cyx_sternheimer_plan_iterate(cyx_stern_plan_t sp, int iters);
/* Synchronises; n_active: columns still iterating.  info, col_iters and
 * col_relres as for cyx_sternheimer_solve (may be NULL). */
int cyx_sternheimer_plan_status(cyx_stern_plan_t sp, int* n_active, cyx_stern_info* info, int32_t* col_iters,
                                double* col_relres);
int cyx_sternheimer_plan_destroy(cyx_stern_plan_t sp);

/* Right-hand sides of the Sternheimer equations of one perturbation
 * (src/phonon.m:1358-1444, "Method-2", q = 0):
 *   dH psi_n = dVeff .* psi_n + sum_J gamma_J (chi_J dchi_J^H + dchi_J chi_J^H) W psi_n
 *   G = psi^H W dH psi                       (psi_DH_psi, nv x nv)
 *   B_n = sum_m psi_m G_mn coeff_mn - dH psi_n
 *   dlambda_n = G_nn                         (S.Dlambda_dR)
 * `dham` is a handle of the same grid and channel whose potential is dVeff
 * (S.Dveff_dR(:,scf_natmc)) and whose projectors carry the projector
 * derivatives as the pairs (chi +- dchi) / sqrt(2) with gammas +-gamma_J
 * (their KB term is exactly the first-order nonlocal term above); its
 * kinetic stencil is not used.  psi: device, nv states (ldpsi); B: device,
 * nv columns (ldb); coeff: host, nv x nv column-major, the reference's
 * S.coeff_Pv_psi_kq(1:nv,1:nv,kpt) after its transpose (src/phonon.m:112-158);
 * dlambda: host, nv (may be NULL). */
int cyx_sternheimer_rhs(cyx_ham_t dham, int dtype, const void* psi, int64_t ldpsi, int nv, const double* coeff,
                        void* B, int64_t ldb, double* dlambda);

/* ------------------------------------------------------------ Poisson */
typedef struct cyx_aar_info {
  int32_t iterations; /* count-1 at exit, as src/aar.m:68-72 reports */
  int32_t converged;  /* relres <= tol */
  double relres;      /* ||f - A x|| / ||f|| of the returned x */
  int32_t precond;    /* 0 Jacobi; 1 learned (cyx_poisson_set_preconditioner);
                         2 learned, diverged, solved again with Jacobi (iterations
                         then count the Jacobi solve); 3 exact Lap_std^-1
                         (cyx_poisson_set_spectral) */
  int32_t pinv_eig;   /* Anderson steps whose DF'DF was not certified full rank
                         (lambda_min > 100 m trace eps) and took the
                         eigen-decomposition pseudo-inverse; the others solve
                         by Cholesky, which equals pinv there */
} cyx_aar_info;

/* Device operator for the Poisson problem: S.Lap_std at k = 0
 * (src/initialization.m:734-741) with a Jacobi preconditioner (the diagonal
 * of Lap_std; extras/AAR_old.m:45-46,73-75). */
int cyx_poisson_create(cyx_ctx_t ctx, const cyx_grid_desc* grid, cyx_poisson_t* out);
int cyx_poisson_destroy(cyx_poisson_t ps);

/* x = aar(Lap_std, f, x_guess, tol, max_iter, omega, beta, m, p, Jacobi).
 * Replaces aar (src/aar.m:1-72).  f, x device FP64 length N; x holds the guess
 * on entry (pass use_guess = 0 for the reference's ones(N,1) start) and the
 * solution on exit.  Returns CYX_ERR_NOCONV (with x and info filled) when
 * max_iter is reached. */
int cyx_aar(cyx_poisson_t ps, const double* f, double* x, int use_guess, double tol, int max_iter, double omega,
            double beta, int m, int p, cyx_aar_info* info);

/* S.phi = poissonSolve(S, tol, Isguess) for Cyclix (src/poissonSolve.m:17-51):
 * f = -4 pi (rho + b), then cyx_aar with omega = beta = 0.6, m = 7, p = 6.
 * rho, b, phi device FP64 length N. */
int cyx_poisson_solve(cyx_poisson_t ps, const double* rho, const double* b, double* phi, int is_guess, double tol,
                      int max_iter, cyx_aar_info* info);

/* y = Lap_std x (device FP64, length N): the operator AAR solves with. */
int cyx_poisson_apply(cyx_poisson_t ps, const double* x, double* y);

/* on != 0: every later cyx_aar / cyx_poisson_solve on ps preconditions with
 * the exact inverse of Lap_std in place of the Jacobi diagonal; replaces the
 * ILU factors S.LapPreconL/U that the reference passes to aar
 * (src/poissonSolve.m:36, src/initialization.m:752).  Lap_std's theta and z
 * factors are circulant, so a theta/z Fourier transform splits it into one
 * band matrix per mode, factored once here (host LU, FP64 device factors of
 * nr x 13 x nt x nz doubles).  AAR (src/aar.m, the same iteration and
 * stopping test) then reaches the solution at its first Anderson step:
 * p + 1 iterations.  info->precond reads 3.  It takes precedence over a
 * learned preconditioner set with cyx_poisson_set_preconditioner.  Fails if
 * a band pivot is not negative (Lap_std not negative definite on the grid). */
int cyx_poisson_set_spectral(cyx_poisson_t ps, int on);
/* z = Lap_std^-1 r exactly for ncol real FP64 columns (leading dimension
 * ld >= N), device memory; builds the factors on first use. */
int cyx_poisson_spectral_apply(cyx_poisson_t ps, const double* r, double* z, int ncol, int64_t ld);

/* ------------------------------------------------------------ learned Poisson preconditioner */
/* A linear three-level 3-D convolutional U-Net U on the handle grid (nr, nt,
 * nz divisible by 4): BF16 weights, BF16 matrix-core products, FP32
 * accumulation.  With it, every Anderson iteration of AAR on a Poisson
 * handle is followed by a learned correction x += delta d with
 * d = D^{-1} r + U(D^{-1} r) (r = f - Lap_std x, D the Jacobi diagonal) and
 * delta = <Lap d, r> / <Lap d, Lap d>, the step that minimises the new
 * residual norm; the Jacobi-AAR iteration itself and its stopping test on the
 * true residual are unchanged (src/aar.m, src/poissonSolve.m:36, where the
 * reference uses ILU factors instead).  It has no counterpart in the
 * reference.  Parameters: one FP32 vector (cyx_unet_num_params entries: the
 * layer weights, then a per-radius output scale). */
int cyx_unet_create(cyx_ctx_t ctx, const cyx_grid_desc* grid, cyx_unet_t* out);
int cyx_unet_destroy(cyx_unet_t net);
int64_t cyx_unet_num_params(cyx_unet_t net);
/* host FP32 parameters in/out; reset_optimizer zeroes the Adam state */
int cyx_unet_set_params(cyx_unet_t net, const float* params, int reset_optimizer);
int cyx_unet_get_params(cyx_unet_t net, float* params);
/* u = U(x) for nb samples (device FP32, nb x N each) */
int cyx_unet_forward(cyx_unet_t net, const float* x, float* u, int nb);
/* loss = (1/nb) sum_b ||U(x_b) - t_b||^2 / ||t_b||^2 and its gradient with
 * respect to the parameters (host FP32, cyx_unet_num_params entries) */
int cyx_unet_gradient(cyx_unet_t net, const float* x, const float* target, int nb, float* grad, double* loss);
/* one Adam step (beta 0.9 / 0.999) on that loss; loss before the step */
int cyx_unet_train_step(cyx_unet_t net, const float* x, const float* target, int nb, float lr, double* loss);
/* x_b = scale_b * r_b / D (the network's input for residuals r, FP64 device,
 * nb x N; scale: device FP32 nb values, or NULL for 1) */
int cyx_unet_input(cyx_unet_t net, const cyx_poisson_t ps, const double* r, const float* scale, int nb, float* x);
/* use net (NULL: Jacobi only) in every later cyx_aar / cyx_poisson_solve on ps;
 * net must live on the same context and grid */
int cyx_poisson_set_preconditioner(cyx_poisson_t ps, cyx_unet_t net);

#ifdef __cplusplus
}
#endif
#endif /* CYCLIX_MSPARC_H */
