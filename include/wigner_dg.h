/*
 * wigner_dg.h -- C ABI of the MI355X Wigner transport library (libwdg.so).
 *
 * Drop-in boundary for the stationary 1D-x x 1D-k Wigner solve of the
 * reference (jonasneu1207/real_space_DG, MATLAB).  The reference's host code
 * reaches its operator only through a small plug-in surface:
 *
 *   [A,rhs]  = get_SysM(mat,p,EM)                   DG_code/core/get_SysM.m:1
 *   sysInfo.apply(u) / assembleRhs() / getDiagonal() DG_code/core/get_SysM_full2D.m:80-86
 *   pre.apply(r)                                    DG_code/core/get_LineTransversePreconditioner_full2D.m:113
 *   gmres(A*M^-1, b, restart, tol, maxit)           DG_code/solve_transport_DG_full2D.m:462-484
 *   rho_u = phi*rho ; n ; j                          DG_code/solve_transport_DG_2D.m:114-156
 *
 * Each entry point below replaces one of those and cites it.  All vectors
 * are interleaved complex double (re,im) in the reference's permuted index
 * order ((e*N_K_chi + a)*Nk + j), k fastest (DG_code/core/get_Diff.m:242-259).
 * Host buffers are owned by the caller (read-only inputs, write-only
 * outputs); the context owns every device allocation.  One HIP stream per
 * context; every call is synchronous at its return.  Functions return
 * WDG_OK (0) or a negative status; wdg_last_error() gives the message.
 */
#ifndef WIGNER_DG_H
#define WIGNER_DG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The three the reference raises itself carry its error
 * identifiers; the others are this library's own (the MATLAB shim raises
 * them as wdg:<Name>). */
enum {
  WDG_OK = 0,
  WDG_ERR_ARG = -1,            /* wdg:InvalidArgument                              */
  WDG_ERR_HIP = -2,            /* wdg:DeviceError, a HIP runtime failure           */
  WDG_ERR_MEMORY = -3,         /* wdg:OutOfMemory, device (or host) memory exhausted */
  WDG_ERR_STATE = -4,          /* wdg:CallOrder, e.g. no potential set yet         */
  WDG_ERR_NUMERIC = -5,        /* wdg:NumericalFailure, singular factor / non-finite value */
  WDG_ERR_UNSUPPORTED = -6,    /* wdg:Unsupported, option outside the restated method */
  WDG_ERR_INITIAL_RHO = -7,    /* DG:Full2D:InvalidInitialRho, a non-finite warm start
                                  (DG_code/solve_transport_DG_full2D.m:508-525; the size
                                  check is the caller's: numel(rho0) == wdg_size)    */
  WDG_ERR_GMRES_MEMORY = -8,   /* DG:Full2D:GmresMemoryLimit (solve_transport_DG_full2D.m:183-203) */
  WDG_ERR_LINE_MEMORY = -9     /* DG:Full2D:LineMemoryLimit
                                  (DG_code/core/get_LineTransversePreconditioner_full2D.m:33-43) */
};

enum { WDG_FERMI_MSA_3D = 0,      /* pure 1D Fermi-Dirac, DG_code/core/verteilung.m:77-114        */
       WDG_FERMI_MSA_2D = 1,      /* kz-integrated Fermi-Dirac, verteilung.m:118-168              */
       WDG_FERMI_TSU_ESAKI = 2 }; /* both transverse directions integrated: m kT/(pi hbar^2)
                                     ln(1+exp((EF-E)/kT)), the form of verteilung.m:64-65 with the
                                     device's EF, band edge and mass (volume density, m^-3)        */
enum { WDG_PREC_FP64 = 0, WDG_PREC_MIXED = 1 };

/* Geometry: the fields of mat / mat.dg.params that initParams reads
 * (DG_code/initParams.m:58-62,110-111,155-191). */
typedef struct wdg_desc {
  int32_t n_chi;       /* DG elements in chi: (mat.Nx-1)/(N_K_chi-1)          */
  int32_t n_k_chi;     /* nodes per element N_K_chi (degree + 1), 2..8        */
  int32_t n_xi;        /* FV cells in xi, mat.dg.params.N_xi (even)          */
  int32_t nk;          /* plane waves, mat.dg.params.N (must equal n_xi)     */
  double  l_chi;       /* device length [m] = mat.L_x*1e-9                    */
  double  l_xi;        /* coherence length [m] = mat.dg.params.Ly             */
} wdg_desc;

/* Potential-dependent data: get_SysM(mat,p,EM) inputs
 * (DG_code/solve_transport_DG_2D.m:44-49,91; verteilung.m:1-170). */
typedef struct wdg_potential {
  int32_t n_mat;        /* material grid points mat.Nx = n_chi*(n_k_chi-1)+1 */
  int32_t fermi_model;  /* WDG_FERMI_*                                       */
  const double* x_mat;  /* material grid [m] (mat.x*1e-9), ascending         */
  const double* em;     /* subband energy EM on x_mat [eV]                   */
  double m_left;        /* x-mass at the first material point [kg]           */
  double m_right;       /* x-mass at the last material point [kg]            */
  double ef_left;       /* EfL [eV]                                          */
  double ef_right;      /* EfR [eV]                                          */
  double temp;          /* mat.Temp [K]                                      */
  double deg_factor;    /* valley degeneracy deg_factor(IV)                  */
  double me_z;          /* mat.me_z_ch (DG_MSA_2D only)                      */
} wdg_potential;

typedef struct wdg_solve_opts {
  double  tol;          /* true relative residual target (full2D_solverTol)        */
  int32_t restart;      /* GMRES restart m (full2D_gmresRestart)                   */
  int32_t max_iter;     /* cap on total inner iterations                           */
  int32_t precision;    /* WDG_PREC_FP64 or WDG_PREC_MIXED                         */
  double  inner_tol;    /* mixed: relative tolerance of each FP32 inner solve      */
  int32_t max_outer;    /* mixed: cap on FP64 refinement steps                     */
  int32_t use_precond;  /* 1: right line preconditioner, 0: none                   */
  int32_t use_graph;    /* 1: replay each restart cycle from a captured HIP graph   */
  int32_t orthogonalization; /* 0: CGS with DGKS second pass, 1: single-pass CGS      */
  int32_t restart_max;  /* adaptive restart cap (<= restart: fixed restart)           */
  int32_t warm_device;  /* 1 and x0 == NULL: start from the context's last solution
                           (kept on the device across wdg_set_potential calls)     */
  int32_t krylov_basis; /* mixed: storage of the inner Krylov basis, WDG_BASIS_WORKING
                           (FP32), WDG_BASIS_BF16 or WDG_BASIS_FP16 (scaled by 256)  */
} wdg_solve_opts;

enum { WDG_BASIS_WORKING = 0, WDG_BASIS_BF16 = 1, WDG_BASIS_FP16 = 2 };

typedef struct wdg_solve_info {
  int32_t flag;                 /* 0 converged (true residual <= tol), 1 not    */
  int32_t iterations;           /* inner GMRES iterations (all refinement steps) */
  int32_t outer_iterations;     /* FP64 refinement steps (1 for the FP64 path)   */
  int32_t restarts;             /* restart cycles run                            */
  double  true_relres;          /* ||b - A x|| / ||b|| in FP64                   */
  double  reported_relres;      /* GMRES's own (Givens) estimate at exit         */
  double  initial_true_relres;  /* of x0 (1 for the zero start)                  */
  double  seconds;              /* device wall time of the solve                 */
  int32_t reorth_steps;         /* Arnoldi steps that took a second CGS pass     */
} wdg_solve_info;

typedef struct wdg_ctx wdg_ctx;

/* Message of the last failing call on ctx (or of the last failed create
 * when ctx is NULL).  Never NULL. */
const char* wdg_last_error(const wdg_ctx* ctx);

/* Grid, basis phi, DG element matrices; device buffers for one instance.
 * Replaces initParams + Basisfunktionen_Q_m (DG_code/initParams.m:58-221,
 * DG_code/core/Basisfunktionen_Q_m.m:3-33). */
int wdg_create(const wdg_desc* desc, wdg_ctx** out);
int wdg_destroy(wdg_ctx* ctx);

/* Number of unknowns N = n_chi*n_k_chi*nk. */
int64_t wdg_size(const wdg_ctx* ctx);

/* Build Theta[V]+CAP tables, inflow data and the line preconditioner for a
 * potential.  Replaces get_SysM(mat,p,EM) (DG_code/core/get_SysM.m:1-16)
 * and the inflow part of initParams (verteilung, DG_code/initParams.m:220). */
int wdg_set_potential(wdg_ctx* ctx, const wdg_potential* pot);

/* rhs of get_Diff (DG_code/core/get_Diff.m:253-259); sysInfo.assembleRhs. */
int wdg_rhs(wdg_ctx* ctx, double* rhs);

/* y = A u with A = A_diff + G (sysInfo.apply, DG_code/core/get_SysM_full2D.m:124-127
 * for the 1D operator of get_SysM.m:3-8).  precision WDG_PREC_FP64 gives the
 * FP64 operator; WDG_PREC_MIXED gives the inner operator of the mixed
 * solver (FP32 transport, split-BF16 MFMA Theta with FP32 accumulation). */
int wdg_apply(wdg_ctx* ctx, const double* u, double* y, int32_t precision);

/* One part of the operator: WDG_PART_TRANSPORT gives y = A_diff u
 * (diff.apply), WDG_PART_POTENTIAL gives y = G u, Theta[V] with the CAP
 * (drift.apply), WDG_PART_ALL their sum as wdg_apply: the two handles that
 * sysInfo.apply adds (DG_code/core/get_SysM_full2D.m:124-127; for the 1D
 * operator A = A_diff + G of DG_code/core/get_SysM.m:3-8).  1D contexts
 * only (full-2D: WDG_PART_ALL). */
enum { WDG_PART_TRANSPORT = 1, WDG_PART_POTENTIAL = 2, WDG_PART_ALL = 3 };
int wdg_apply_parts(wdg_ctx* ctx, const double* u, double* y, int32_t precision, int32_t parts);

/* z = M^-1 r for the line preconditioner (pre.apply,
 * DG_code/core/get_LineTransversePreconditioner_full2D.m:199-239, restated
 * for the k basis: exact per-k block-bidiagonal upwind solve along x of
 * A_diff + diag_k(G)). */
int wdg_precond(wdg_ctx* ctx, const double* r, double* z);

/* diag(A) (sysInfo.getDiagonal, DG_code/core/get_SysM_full2D.m:83). */
int wdg_diagonal(wdg_ctx* ctx, double* d);

/* GMRES solve of A x = rhs (solveStationarySystem gmres branch,
 * DG_code/solve_transport_DG_full2D.m:276-334 with runGmresWithProgress
 * :462-484): RHS normalisation, right preconditioning, warm start x0 as a
 * correction with the absolute target tol*||b||, true residual check that
 * overrides the flag.  x0 may be NULL (zero start); x may be NULL (the
 * solution stays on the device for wdg_observables). */
int wdg_solve(wdg_ctx* ctx, const wdg_solve_opts* opts, const double* x0, double* x,
              wdg_solve_info* info);

/* The context's last solution (of wdg_solve, wdg_solve_batch or a transient
 * run), N interleaved complex doubles in physical units: the rho that
 * solve_transport_DG_full2D returns (DG_code/solve_transport_DG_full2D.m:67-129)
 * and solve_transport_DG_2D with st_init (DG_code/solve_transport_DG_2D.m:211-225),
 * for solutions left on the device.  WDG_ERR_STATE before any solve. */
int wdg_solution(wdg_ctx* ctx, double* x);

/* Residual history of the context's last solve (wdg_solve, wdg_solve_batch,
 * or the last step of a Crank-Nicolson run): the resvec of the reference's
 * GMRES call on the normalised system (DG_code/solve_transport_DG_full2D.m:306-308):
 * [0] the initial residual, [k] the GMRES estimate after Arnoldi step k, both
 * relative to ||rhs||.  The reference reports resvec * ||rhs|| as
 * info.solve.resvec (:324,333); multiply by ||rhs|| for that.  *n receives
 * the full length (iterations + 1, 0 before any solve; at most
 * min(max_iter + 1, 2^22)); min(n, cap) values are copied to out. */
int wdg_solve_resvec(wdg_ctx* ctx, double* out, int32_t cap, int32_t* n);

/* wdg_solve (zero start, solution left on the device) on n independent
 * instances together: every instance runs its restart cycles on its own
 * stream and the host polls each one once per cycle, so all unconverged
 * instances are in flight at once.  Replaces the serial loop over
 * transverse modes and valleys of DG_code/solve_transport_DG_2D.m:62-209
 * (one stationary solve per mode).  Instances may differ in grid and
 * potential, 1D and full-2D alike; infos has n entries.  On failure the
 * message of wdg_last_error(NULL) names the instance. */
int wdg_solve_batch(wdg_ctx* const* ctxs, int32_t n, const wdg_solve_opts* opts, wdg_solve_info* infos);

/* Observables of rho (or of the last solution when rho is NULL):
 * rho_u = phi*rho per node, n = Re rho_u at xi=0, j = q*hbar/m*Im d rho_u/dxi
 * (DG_code/solve_transport_DG_2D.m:114-156).  m_nodes: x-mass per DG node
 * [kg] (n_chi*n_k_chi).  n, j: n_chi*n_k_chi doubles.  rho_u may be NULL,
 * otherwise n_xi x (n_chi*n_k_chi) interleaved complex, column-major
 * (xi fastest) like RHO_DG's transpose. */
int wdg_observables(wdg_ctx* ctx, const double* rho, const double* m_nodes, double* n,
                    double* j, double* rho_u);

/* ---- self-consistent electrostatics (RTD Gummel loop) ---- */
/* 1D Poisson problem on the material grid, Dirichlet contacts phi(0) = 0,
 * phi(L) = bias (RTD/eval_Poisson2D_rtd.m, RTD/solve_Poisson2D_rtd.m with
 * Ny = 1).  Electron energy is vc - phi [eV]. */
typedef struct wdg_poisson_desc {
  int32_t n_mat;          /* material grid points (mat.Nx), uniform spacing  */
  int32_t method;         /* 0 'direct', 1 'newton-raphson' (Boltzmann Jacobian) */
  const double* x_mat;    /* material grid [m]                                */
  const double* eps_r;    /* relative permittivity per node                   */
  const double* net_doping; /* Nd - Na per node [m^-3]                        */
  double temp;            /* [K]                                              */
} wdg_poisson_desc;

typedef struct wdg_poisson_info {
  double physical_relres_before;  /* ||F_free||_inf / scale at phi_in        */
  double unmixed_step;            /* ||delta||_inf [V]                        */
  double charge_integral;         /* integral of the projected density [m^-2] */
  double dg_integral;             /* integral of the DG density [m^-2]        */
} wdg_poisson_info;

/* Geometry of the Poisson problem; builds the conservative DG -> dual-volume
 * density projection (RTD/project_DG_density_RTD.m) and the operator. */
int wdg_poisson_setup(wdg_ctx* ctx, const wdg_poisson_desc* desc);

/* One FP64 Poisson update on the device from the density of the context's
 * last solution: phi_out = phi_in + delta (unmixed candidate; Gummel mixing
 * is the caller's).  n_out (may be NULL) receives the projected density.
 * RTD/make_RTD_poisson_step.m:44-54 + RTD/solve_Poisson2D_rtd.m:85-130. */
int wdg_poisson_step(wdg_ctx* ctx, const double* phi_in, double bias, double* phi_out, double* n_out,
                     wdg_poisson_info* info);

/* ---- device-resident variants (for callers that keep data in HBM) ---- */
/* u, y: device pointers, interleaved complex of the precision's type
 * (double for FP64, float for mixed).  Enqueued on the context stream and
 * synchronised before return. */
int wdg_apply_device(wdg_ctx* ctx, const void* d_u, void* d_y, int32_t precision);
int wdg_precond_device(wdg_ctx* ctx, const void* d_r, void* d_z, int32_t precision);

/* Average device time of one operator apply (wdg_apply_device's kernels) in
 * the given precision: nrep applies enqueued back to back on the context
 * stream between one HIP event pair, on the context's own work buffers (the
 * apply's cost does not depend on the values).  A measurement helper with
 * no counterpart in the reference: per-kernel event pairs around single
 * launches also time the host's submission gaps. */
int wdg_time_apply(wdg_ctx* ctx, int32_t precision, int32_t nrep, double* ms_per_apply);

/* wdg_precond_device that allocates nothing: for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  1D contexts
 * only.  The chunk-map workspace of the line sweeps (n_chunks * nk * 2
 * complex doubles) lives in a file-scope device array of 8 MiB that each
 * GPU the library runs on holds until the process ends; grids that need
 * more return WDG_ERR_ARG.  All calls on one GPU share that array: calls
 * that overlap on one GPU (several streams or threads) race.  Results equal
 * wdg_precond_device's bit for bit. */
int wdg_precond_device_noalloc(wdg_ctx* ctx, const void* d_r, void* d_z, int32_t precision);

/* ---- host-only helpers (no GPU needed) ---- */
/* The host-side setup of wdg_create + wdg_set_potential without a device:
 * rhs (2N doubles) and the per-element transport matrices K1,K2L,K2R,K3,K5
 * (5*n*n doubles, row-major, xi-scaled inverse mass, get_Diff.m:198-221).
 * Either output may be NULL. */
int wdg_host_setup(const wdg_desc* desc, const wdg_potential* pot, double* rhs, double* kmats);

/* Explicit RK4 time stepping of d rho/dt = b_scale*rhs - a_scale*A rho
 * (DG_code/core/transient.m:12-16; DG_code/solve_transport_DG_2D_transient_SC_new.m:145-158,
 * where A is prescaled by 2e-10 and rhs is not).  Starts from rho0 (host,
 * 2N doubles) or, when rho0 is NULL, from the context's last solution;
 * the final state replaces that solution (rho_out may be NULL).  Every
 * save_every steps the observables of the state (as wdg_observables, with
 * m_nodes) are written to column s of n_t and j_t (n_chi*n_k_chi x n_saves,
 * column-major); n_saves = nsteps / save_every must not exceed max_saves.
 * The potential and inflow of the last wdg_set_potential stay fixed. */
typedef struct wdg_transient_opts {
  double  dt;          /* time step [s] (mat.dt)                                 */
  int32_t nsteps;      /* number of steps (mat.Nt)                               */
  int32_t save_every;  /* observables every save_every steps (>= 1)              */
  double  a_scale;     /* factor on A (the reference uses 2e-10)                 */
  double  b_scale;     /* factor on rhs (the reference uses 1)                   */
} wdg_transient_opts;
int wdg_transient_rk4(wdg_ctx* ctx, const wdg_transient_opts* opts, const double* rho0, double* rho_out,
                      const double* m_nodes, double* n_t, double* j_t, int32_t max_saves);

/* Crank-Nicolson (trapezoidal) time stepping of the same ODE,
 * (I + dt/2 a_scale A) rho+ = (I - dt/2 a_scale A) rho + dt b_scale rhs,
 * with one device GMRES solve per step warm-started from rho (tolerance tol
 * relative to the step's right-hand side, precision / krylov_basis as in
 * wdg_solve_opts).  No counterpart in the reference, whose full-2D transient
 * entry points are stubs (DG_code/solve_transport_DG_full2D_transient_flatband.m:8-11):
 * the implicit step that lets the time step follow the physics instead of
 * the operator's stiffness.  Starts from rho0 (host) or, when NULL, from the
 * context's last solution; the final state becomes that solution (read it
 * with wdg_observables / wdg2_observables, rho = NULL).  Works on 1D and
 * full-2D contexts; full-2D Jacobi / row-sum preconditioners are rebuilt for
 * I + dt/2 a_scale A (use_precond), the 1D line and the line-transverse
 * preconditioners are not used for the shifted system. */
/* use_graph is ignored (cycles are launched kernel by kernel) for FP64 runs
 * on full-2D contexts: at 33.5M unknowns such a run with graph replay ended
 * in a host-side fault inside the HIP runtime whose cause is not identified. */
typedef struct wdg_cn_opts {
  double  dt;            /* time step [s]                                   */
  int32_t nsteps;
  double  a_scale, b_scale;
  double  tol;           /* per-step GMRES tolerance                        */
  int32_t restart, max_iter, precision, krylov_basis, use_precond;
  int32_t use_graph;     /* 1: replay the restart cycles from captured HIP graphs */
} wdg_cn_opts;
typedef struct wdg_cn_info {
  int32_t steps;               /* steps taken                               */
  int32_t iterations;          /* GMRES iterations, all steps               */
  int32_t max_step_iterations;
  int32_t flag;                /* 0: every step converged                   */
  double  max_relres;          /* largest per-step true relative residual   */
  double  seconds;             /* device time of the run                    */
} wdg_cn_info;
int wdg_transient_cn(wdg_ctx* ctx, const wdg_cn_opts* opts, const double* rho0, double* rho_out,
                     wdg_cn_info* info);
/* wdg_transient_cn on n independent instances together: each step launches
 * every instance's right-hand side, then runs all the step solves with their
 * restart cycles interleaved on the instances' own streams (as
 * wdg_solve_batch), so the GPU keeps every unconverged instance in flight.
 * rho0 / rho_out are arrays of n host pointers (the arrays themselves may be
 * NULL: all instances start from their last solution / no copy back); infos
 * has n entries.  All instances take the same options.  No counterpart in the
 * reference, which has no transient for these instances (its full-2D
 * transient entry points are stubs); on failure the message of
 * wdg_last_error(NULL) names the instance. */
int wdg_transient_cn_batch(wdg_ctx* const* ctxs, int32_t n, const wdg_cn_opts* opts, const double* const* rho0,
                           double* const* rho_out, wdg_cn_info* infos);

/* Formulation of the Theta[V] apply.  WDG_THETA_FFT evaluates phi U and
 * i phi^H Z as chirp-z transforms (FFTs of length 2^p 3^q >= nk+n_xi-1) in
 * one fused kernel per element; WDG_THETA_GEMM runs them as dense batched
 * MFMA GEMMs (split-BF16 in the mixed inner solver, FP64 otherwise).  Both
 * compute the same operator as get_Drift_accel_trial_permutedShell.m:95-151
 * (G_q = i phi^H C_q phi); AUTO (the default) picks FFT when the grid allows
 * it.  No counterpart in the reference: an implementation choice.
 * Returns WDG_ERR_ARG if FFT is requested for a grid it does not apply to. */
enum { WDG_THETA_AUTO = 0, WDG_THETA_FFT = 1, WDG_THETA_GEMM = 2 };
int wdg_set_theta_method(wdg_ctx* ctx, int32_t method);
/* The formulation in use (WDG_THETA_FFT or WDG_THETA_GEMM), -1 for NULL. */
int wdg_theta_method(const wdg_ctx* ctx);

/* Memory guards of the reference.  gmres_max_basis_bytes bounds the
 * estimated Krylov basis 16 * N * (restart + 1) bytes of every solve
 * (full2D_gmresMaxBasisBytes, 9 GiB on the GPU for the line-transverse
 * preconditioner, solve_transport_DG_full2D.m:183-203; 0 or +inf: no limit,
 * the default); a solve above it returns WDG_ERR_GMRES_MEMORY.
 * line_max_factor_bytes bounds the line-transverse factors of
 * wdg2_set_line_precond, 32 nT^2 nModes + 32 nLocal^2 nModes nElem +
 * 16 nModes^2 + 8 nT nModes bytes (full2D_lineMaxFactorBytes, 4 GiB by
 * default, get_LineTransversePreconditioner_full2D.m:33-43); above it, or
 * when it is not positive and finite, the call returns WDG_ERR_LINE_MEMORY. */
int wdg_set_memory_limits(wdg_ctx* ctx, double gmres_max_basis_bytes, double line_max_factor_bytes);

/* Kernel of the FP32 (mixed inner) FFT Theta apply:
 * WDG_FFT_BLOCK_256 runs the chirp-z kernel with 256 threads that loop over
 * the element's nodes; WDG_FFT_BLOCK_WIDE runs it with ~1024 threads, four
 * or more transforms at once, for n_k_chi <= 5; WDG_FFT_BLOCK_NDFT runs the
 * near-DFT kernel (square grids n_xi = nk = 64 * 2^a 3^b, n_k_chi <= 5):
 * transforms of length nk instead of >= 2 nk - 1, the residual phase
 * exp(i s ubar vbar), |s| = pi/(2 nk), by a Taylor series truncated below
 * 5e-6 relative (two terms for nk >= 512).  AUTO (the default) runs the
 * chirp-z kernel in the shape an autotuner picks: before the first mixed
 * solve of the context it times both workgroup shapes on the context's grid
 * (cached per grid shape for the process; until then the shape that keeps
 * more waves resident per CU).  NDFT is never chosen
 * by AUTO (its truncation, ~3e-7 relative at nk = 1024, costs the mixed
 * solver more refinement cycles than its shorter transforms save).  All
 * compute the operator of
 * get_Drift_accel_trial_permutedShell.m:95-151 to FP32 accuracy.  No
 * counterpart in the reference: an implementation choice, exposed so that
 * every variant can be tested.  Returns WDG_ERR_ARG when the grid has no
 * FFT plan or the requested kernel does not fit it.  wdg_fft_block gives
 * the threads per workgroup in use (-1: no FFT plan). */
enum { WDG_FFT_BLOCK_AUTO = 0, WDG_FFT_BLOCK_256 = 1, WDG_FFT_BLOCK_WIDE = 2, WDG_FFT_BLOCK_NDFT = 3 };
int wdg_set_fft_block(wdg_ctx* ctx, int32_t mode);
int wdg_fft_block(const wdg_ctx* ctx);

/* Enable per-kernel timing with HIP events on the context stream for the
 * kernels selected by mask (bit i = entry i of wdg_profile's list); 0 turns
 * it off.  Resets the accumulated statistics. */
int wdg_set_profiling(wdg_ctx* ctx, uint32_t mask);

/* Per-kernel statistics collected while profiling is enabled: average
 * device time [ms] per launch, launches, and the total algorithmic work
 * (bytes for memory-bound kernels, flops for the Theta GEMMs).  names gets
 * the comma-separated kernel list; each array holds max_entries values. */
int wdg_profile(wdg_ctx* ctx, char* names, int32_t names_len, double* avg_ms, int64_t* launches,
                double* units, int32_t max_entries, int32_t* n_entries);

/* ==========================================================================
 * Full-2D rho-basis operator F(rho_x, rho_y, X, Y)
 * (DG_code/core/get_SysM_full2D.m:1-91, DG_code/solve_transport_DG_full2D.m)
 *
 * Vector order as initParams_full2D.m:4-8,105-107: rho_x fastest, then
 * rho_y, the X DG DOF and the Y DG DOF (N = n_rho_x*n_rho_y*n_X_dof*n_Y_dof
 * interleaved complex doubles).  A full-2D context is a wdg_ctx: wdg_apply,
 * wdg_diagonal, wdg_precond, wdg_solve, wdg_solve_batch, wdg_rhs, wdg_size,
 * the profiling calls and wdg_destroy work on it; the 1D-only calls
 * (wdg_set_potential, wdg_observables, wdg_poisson_*, wdg_transient_rk4,
 * wdg_set_theta_method) return WDG_ERR_UNSUPPORTED.
 * ========================================================================== */

/* Grid (initParams_full2D.m:58-111): relative FV axes and rectangular DG axes. */
typedef struct wdg2_desc {
  int32_t n_rho_x, n_rho_y;   /* FV cells of rho_x, rho_y (N_rho_x, N_rho_y), >= 2           */
  int32_t nl_x, nl_y;         /* DG nodes per element (N_K_X, N_K_Y), 2..4                   */
  int32_t ne_x, ne_y;         /* DG elements (N_el_X, N_el_Y)                                */
  double  delta_x, delta_y;   /* FV cell widths; central difference D = (-1, 0, +1)/(2 delta)
                                 (initParams_full2D.m:199-214)                                */
  int32_t periodic_y;         /* 1: periodic Y (initParams_full2D.m:65-87): the central
                                 difference along rho_y, the banded Y parts and the
                                 relative vertex map wrap around, the potential is
                                 sampled periodically in Y (get_Drift_full2D.m:180-184,
                                 650-666); Y face parts must not be given               */
} wdg2_desc;

/* One operator part coef * kron(C, R) (get_Diff_full2D.m:463-474 addOperatorPart),
 * C acting along centre axis `axis` and R along the relative axis of the same
 * name.  Centre factors:
 *   BANDED - block tridiagonal along the axis: per element e an nl x nl
 *            row-major block (c_diag, ne*nl*nl), the column that couples to
 *            the last node of e-1 (c_lower, ne*nl) and to the first node of
 *            e+1 (c_upper, ne*nl): H1 / interior H2, get_Diff_full2D.m:206-266.
 *   FACE   - lift (c_diag, nl values) of the trace node at the low (face 0,
 *            first element, node 0) or high end (face 1, last element, last
 *            node) into that element; rows of transverse DOF t are kept when
 *            mask[t] != 0 (mask may be NULL): leftBoundary/rightBoundary and
 *            the contact masks, get_Diff_full2D.m:393-452.
 * Relative factors: CENTRAL -i D (Ax1D/Ay1D, initParams_full2D.m:99-102),
 * IDENTITY (Rusanov beta I, get_Diff_full2D.m:454-462), DENSE n_rho x n_rho
 * complex row-major r_dense (|A|, A+ R, A- R: get_Diff_full2D.m:295-347,718-730).
 * BANDED parts pair with any relative factor, FACE parts with DENSE. */
enum { WDG2_AXIS_X = 0, WDG2_AXIS_Y = 1 };
enum { WDG2_CENTER_BANDED = 0, WDG2_CENTER_FACE = 1 };
enum { WDG2_REL_CENTRAL = 0, WDG2_REL_IDENTITY = 1, WDG2_REL_DENSE = 2 };
typedef struct wdg2_part {
  int32_t axis, center, rel, face;
  double  coef;
  const double*  c_diag;
  const double*  c_lower;
  const double*  c_upper;
  const int32_t* mask;
  const double*  r_dense;
} wdg2_part;

/* Creates a full-2D context from the transport parts (get_Diff_full2D.m:80-146).
 * The potential must follow (wdg2_set_potential). */
int wdg2_create(const wdg2_desc* desc, const wdg2_part* parts, int32_t n_parts, wdg_ctx** out);

/* Potential and CAP (get_Drift_full2D.m:1-209, get_CAP_full2D.m:1-39).
 * FV_CONSISTENT: (1/16) P^T diag(DeltaV) P on the relative vertices, projected
 *   over each centre element's tensor quadrature (get_Drift_full2D.m:331-364):
 *   center_x/_y are the quadrature coordinates (ne*nq per axis, element-major),
 *   rel_x/_y the relative vertex coordinates (n_rho+1), interp_* (nq x nl) the
 *   local basis at the quadrature points and proj_* (nq x nl) the projection
 *   w_q (interp M^-1)^T factors of centerProjection.
 * COLLOCATED: diagonal DeltaV at the DG nodes (center_x/_y: nodes, n_dof per
 *   axis) and cell centres (rel_x/_y: n_rho).
 * DeltaV = V(X + rho_x/2, Y + rho_y/2) - V(X - rho_x/2, Y - rho_y/2) with V
 * bilinear on the ascending material grid after clamping each coordinate
 * (evaluateExtendedPotential, get_Drift_full2D.m:650-666).  The operator adds
 * i*pot_coef*DeltaV and cap_coef*(cap_x(rho_x) + cap_y(rho_y)). */
enum { WDG2_POT_FV_CONSISTENT = 0, WDG2_POT_COLLOCATED = 1 };
typedef struct wdg2_potential {
  int32_t mode;
  int32_t nq_x, nq_y;          /* quadrature points per element axis (FV), 1..8       */
  int32_t nx_grid, ny_grid;    /* material grid, >= 2 each                            */
  const double* x_grid;        /* [m], ascending                                     */
  const double* y_grid;
  const double* v;             /* nx_grid*ny_grid [V], x fastest                      */
  const double* center_x;
  const double* center_y;
  const double* rel_x;
  const double* rel_y;
  const double* cap_x;         /* n_rho_x CAP profile (get_CAP.m)                    */
  const double* cap_y;
  const double* interp_x;      /* FV only */
  const double* interp_y;
  const double* proj_x;
  const double* proj_y;
  double pot_coef;             /* drift_scale*Q_drift (get_Drift_full2D.m:161)       */
  double cap_coef;             /* drift_scale*cap_scale*Q_drift (:162)               */
} wdg2_potential;
int wdg2_set_potential(wdg_ctx* ctx, const wdg2_potential* pot);

/* Source of Delta V in the FV-consistent apply.  WDG2_DV_TABLE (the
 * default) stores Delta V per (element, quadrature point, relative vertex)
 * in HBM (12 bytes per entry for both precisions) and streams it;
 * WDG2_DV_INTERP interpolates the material grid inside the apply kernel
 * (the grid and the per-sample cell tables sit in LDS), so an apply streams
 * only u and y but evaluates two bilinear interpolations per sample.  The
 * apply is bound by the latency of its dependent load rounds, not by HBM
 * bytes, and the table form measured 2.5x faster (FET slab, 33.5M
 * unknowns: 0.64 vs 1.60 ms per FP32 apply).  Both compute the same values
 * (the table is filled by the same interpolation); the choice depends on
 * the mode alone.  Takes effect at once (or at wdg2_set_potential).  No
 * counterpart in the reference: an implementation choice. */
enum { WDG2_DV_INTERP = 0, WDG2_DV_TABLE = 1 };
int wdg2_set_dv_mode(wdg_ctx* ctx, int32_t mode);

/* The inflow right-hand side assembled by the caller (sysInfo.assembleRhs(),
 * get_Diff_full2D.m:609-616), N interleaved complex doubles; uploaded once
 * (solve_transport_DG_full2D.m:372). */
int wdg2_set_rhs(wdg_ctx* ctx, const double* rhs);

/* Contact-mode reservoirs and the inflow right-hand side built on the
 * device from host-solved transverse contact modes (the per-bias rebuild of
 * get_ContactModeReservoirRho_full2D.m:35-137 with the longitudinal
 * occupation of :406-424, and the X-face inflow parts of
 * get_Diff_full2D.m:393-452 / assembleRhs :609-616).  Per contact and mode m
 *   f_m(kx) = 2 deg/pi trapz_kz 1/(1 + e^{q(E_m + hb^2 kx^2/2m_x + hb^2 kz^2/2m_z - E_F)/kT})
 *             (n_kz = 0: 2 deg/(1 + e^{...}); exponent clipped to +-80),
 *   F_m(rho_x) = sum_kx cos(rho_x kx) f_m(kx) dkx/(2 pi),
 *   rho(rho_y, rho_x; Y) = sum_m phi_m(Y + rho_y/2) conj(phi_m(Y - rho_y/2)) F_m(rho_x)
 * on the contact rows (mask), and the rhs of the first (source: +A+) and last
 * (drain: -A-) X element, scale_x * lift_a * (A rho), replacing the one
 * wdg2_set_rhs would upload.  The modes, their masses and the Fermi levels
 * (neutrality) stay host work: small dense eigenproblems and a scalar
 * bisection. */
enum { WDG2_MODES_BOX = 0, WDG2_MODES_FOURIER = 1 };
typedef struct wdg2_contact_modes {
  int32_t n_modes;            /* 0..128; 0: no reservoir on this contact          */
  int32_t kind;               /* BOX: phi by linear interpolation on grid, zero
                                 outside; FOURIER: sum_i c_i e^{i k_i (y-y0)}/sqrt(width) */
  int32_t n_grid;             /* grid points (BOX) or Fourier terms (FOURIER)     */
  const double* grid;         /* BOX: ascending Y grid [m] including one zero ghost
                                 node beyond each edge; FOURIER: k_i [1/m]          */
  const double* values;       /* BOX: n_modes x n_grid reals, mode-major;
                                 FOURIER: n_modes x n_grid complex (interleaved)    */
  double y0, width;           /* FOURIER: phase origin and periodic cell width [m] */
  const double* energy;       /* n_modes mode energies [eV]                       */
  const double* mass_x;       /* n_modes longitudinal masses [kg]                 */
  const double* mass_z;       /* n_modes k_z masses [kg]                          */
  double fermi_level;         /* E_F [eV]                                          */
  const int32_t* mask;        /* n_Y_dof: nonzero on contact rows                  */
  const double* char_matrix;  /* n_rho_x^2 complex, row-major: A+ (source), A- (drain) */
  const double* lift;         /* n_l_x: face lift into the first / last X element   */
  double* reservoir_out;      /* NULL, or rho: (n_rho_x n_rho_y) x n_Y_dof complex,
                                 rho_x fastest, Y dof slowest                      */
} wdg2_contact_modes;
typedef struct wdg2_reservoir_grid {
  double temp, deg_factor;    /* [K], valley degeneracy                            */
  int32_t n_kx;               /* longitudinal k grid (makeKGrid) and its spacing   */
  const double* kx;
  double dkx;
  int32_t n_kz;               /* 0: no k_z integration                             */
  const double* kz;           /* ascending, from 0                                 */
  const double* rho_x;        /* n_rho_x relative cell centres [m]                 */
  const double* rho_y;        /* n_rho_y relative cell centres [m]                 */
  const double* y_nodes;      /* n_Y_dof DG node coordinates [m]                   */
  double scale_x;             /* transport scale of the X inflow parts             */
} wdg2_reservoir_grid;
int wdg2_set_contact_rhs(wdg_ctx* ctx, const wdg2_reservoir_grid* grid, const wdg2_contact_modes* source,
                         const wdg2_contact_modes* drain);

/* sysInfo.getRowAbsSum() (get_SysM_full2D.m:165-168): transport row sums of
 * the Kronecker parts plus the drift triangle-inequality bound
 * (get_Drift_full2D.m:490-523); N doubles. */
int wdg2_rowabs(wdg_ctx* ctx, double* s);

/* Diagonal preconditioner of the Krylov solve (buildPreconditioner,
 * solve_transport_DG_full2D.m:788-925,1187-1236): JACOBI inverts diag(A)
 * (entries below 1e-12*max(max|d|,1) get identity action) and falls back to
 * ROWABS when such entries exist; ROWABS divides by the row-absolute sums.
 * *used receives the method in effect.  wdg_solve then applies it as a right
 * preconditioner; NONE disables it. */
enum { WDG2_PRECOND_NONE = 0, WDG2_PRECOND_JACOBI = 1, WDG2_PRECOND_ROWABS = 2, WDG2_PRECOND_LINE = 3 };
int wdg2_set_precond(wdg_ctx* ctx, int32_t kind, int32_t* used);

/* Line-transverse preconditioner (get_LineTransversePreconditioner_full2D.m:1-239):
 * M = (sigma I + Xref)(sigma I + Tref)/sigma in the eigenbasis U of -i D_rhoX,
 * applied as z = U T^-1 X^-1 U^H r / sigma (applyInverse, :199-239):
 *   modes       n_rho_x x n_rho_x complex, row-major, column k = eigenvector k;
 *   forward     per mode, nonzero = X sweep from the first element (speed >= 0);
 *   x_inverse   (I + a K1 + |a| K2)^-1 and x_coupling the neighbour block of
 *               xFactors (:174-197): n_rho_x * ne_x * nl_x^2 doubles, mode-major,
 *               then element, each nl_x x nl_x row-major;
 *   t_inverse   (B_k/sigma + I)^-1 per mode (the dense LU pages of :82-99 as
 *               explicit inverses): n_rho_x * nT^2 complex, row-major,
 *               nT = n_rho_y * n_Y_dof, transverse index t = rho_y + n_rho_y*Ydof.
 * The caller computes the factors on the host (the reference's CPU setup);
 * this call uploads them and selects WDG2_PRECOND_LINE for wdg_solve. */
typedef struct wdg2_line_factors {
  double sigma;
  const double*  modes;
  const int32_t* forward;
  const double*  x_inverse;
  const double*  x_coupling;
  const double*  t_inverse;
} wdg2_line_factors;
int wdg2_set_line_precond(wdg_ctx* ctx, const wdg2_line_factors* factors);

/* Zero-separation observables (get_Observables_full2D.m:24-60): per centre
 * DOF i, out[k*ncen + i] = sum_j w[8k+j] * rho[i*nrel + idx[8k+j]],
 * j < counts[k] <= 8, for k = 0 (density weights), 1 (d/drho_x row),
 * 2 (d/drho_y row); 3*ncen complex outputs.  rho NULL: the context's last
 * solution. */
int wdg2_observables(wdg_ctx* ctx, const double* rho, const int32_t* idx, const double* w, const int32_t* counts,
                     double* out);

/* ---- full-2D electrostatics on the device ----
 * The Poisson problem of the self-consistent full-2D loop on the material
 * grid (RTD/eval_Poisson2D_rtd.m:1-216, RTD/solve_Poisson2D_rtd.m:85-130)
 * and the conservative DG -> dual-volume density projection
 * (RTD/project_DG_density_RTD.m:1-69), so that the loop's density,
 * projection and Newton update stay in HBM between transport solves.
 * Unique material nodes u = ix + nx*iy, iy < nu = ny - periodic_y (the
 * periodic duplicate row is the caller's to expand).  The operator A is
 * M of eval_Poisson2D_rtd.m restricted to the unique rows with the
 * duplicate column folded onto row 0, given as COO over unique nodes
 * (volts, x and y in nm: the factor 1e18 of eps0/dx^2); its coupling
 * must stay within one material column or the neighbouring ones (bandwidth
 * nu in the device's y-fastest order), which the 5-point stencil does. */
typedef struct wdg2_poisson_desc {
  int32_t nx, nu;               /* material nodes along x, unique nodes along y              */
  int32_t n_entries;            /* nonzeros of A                                             */
  const int32_t* a_row;         /* n_entries unique-node rows                                */
  const int32_t* a_col;         /* n_entries unique-node columns                             */
  const double* a_val;          /* n_entries values                                          */
  const int32_t* node_type;     /* nx*nu: 1 free, 2 Dirichlet                                */
  const double* net_doping;     /* nx*nu: Nd - Na [m^-3]                                     */
  double temp;                  /* [K] (Boltzmann Jacobian of the Newton update)            */
  int32_t x_dof, y_dof;         /* centre DOFs of the full-2D contexts along X and Y        */
  const double* proj_x;         /* nx*x_dof row-major: integral of each X basis function
                                   over the dual interval of material node ix               */
  const double* proj_y;         /* nu*y_dof row-major: the same along Y (periodic: the two
                                   endpoint half intervals merged into row 0)                */
  const double* volume;         /* nx*nu dual volumes (index ix + nx*iy)                    */
  int32_t n_weights;            /* zero-separation density weights, <= 8                    */
  const int32_t* w_idx;         /* relative index of each weight (get_Observables_full2D.m) */
  const double* w_val;          /* weight values                                             */
} wdg2_poisson_desc;

/* Builds the Poisson data on the context (a full-2D context; the contexts
 * passed to wdg2_poisson_step must share its grid).  WDG_ERR_ARG for
 * sizes, indices or couplings outside the description above. */
int wdg2_poisson_setup(wdg_ctx* ctx, const wdg2_poisson_desc* desc);

typedef struct wdg2_poisson_info {
  double physical_relres_before; /* ||F_free||_inf / scale before the update               */
  double physical_relres_after;  /* the same at the (mixed) updated potential              */
  double linear_relres;          /* ||J delta + F||_inf / ||F||_inf of the banded solve    */
  double unmixed_step;           /* ||delta||_inf [V]                                       */
} wdg2_poisson_info;

/* One FP64 Poisson update on the device (solve_Poisson2D_rtd.m:85-130):
 * the density is the sum over ctxs[0..n-1] of their last solutions'
 * zero-separation densities (n = 1 for one transport instance, the
 * transverse modes of a device otherwise), projected onto the dual
 * volumes; v = phi_in with the Dirichlet values bc on fixed nodes,
 * F = A v - q (n - net), J = A (newton = 0, 'direct') or
 * A - diag(q^2 n / kT) (newton = 1), delta = -J^-1 F by a banded LU
 * without pivoting (J is an M-matrix up to sign), phi_out = v + alpha delta
 * on free nodes.  phi_in, bc, phi_out, cand_out (v + delta, may be NULL)
 * and n_out (projected density [m^-3], may be NULL) hold nx*nu values.
 * The Poisson data are ctxs[0]'s (wdg2_poisson_setup). */
int wdg2_poisson_step(wdg_ctx* const* ctxs, int32_t n, const double* phi_in, const double* bc, int32_t newton,
                      double alpha, double* phi_out, double* cand_out, double* n_out, wdg2_poisson_info* info);

#ifdef __cplusplus
}
#endif
#endif /* WIGNER_DG_H */
