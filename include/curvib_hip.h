/*
 * curvib_hip.h -- C ABI of the MI355X (gfx950) fractional-step core of
 * CURVIB.  Plain pointers and sizes only; every entry point returns an int
 * status (CV_OK = 0).
 *
 * Array layout at this boundary is the reference's PETSc DMDA global layout:
 * storage extents mx = IM+1, my = JM+1, mz = KM+1 (init.c:161-165); scalar
 * fields are mx*my*mz doubles, Cmpnts fields 3*mx*my*mz doubles interleaved
 * (x,y,z) (variables.h:17-19); natural order i + mx*(j + my*k)
 * (poisson.c:95).  Host-pointer entry points are synchronous: they copy in,
 * run on the context's HIP stream and copy out.  All device state (metrics,
 * operator coefficients, multigrid hierarchy, Krylov space, flow fields) is
 * allocated once by cv_ctx_create/cv_set_geometry and owned by the context.
 *
 * Threading: calls on one context are serialised by the caller; distinct
 * contexts may be driven from distinct host threads.
 */
#ifndef CURVIB_HIP_H
#define CURVIB_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference returns PetscErrorCode and never checks
 * convergence; poisson.c:1828-1983) */
#define CV_OK                 0
#define CV_ERR_ARG           -1
#define CV_ERR_OOM           -2
#define CV_ERR_NOT_CONVERGED -3
#define CV_ERR_NONFINITE     -4
#define CV_ERR_HIP           -5
#define CV_ERR_STATE         -6

/* Poisson preconditioner precision */
#define CV_FP64   0   /* FP64 V-cycle */
#define CV_MIXED  1   /* FP32 V-cycle inside the FP64 FGMRES */
#define CV_MIXED16 2  /* as CV_MIXED, the V-cycle's stencil reading its face
                         coefficients in FP16 (scaled by a power of two per
                         level); FGMRES still solves the FP64 operator */

/* multigrid smoothers */
#define CV_SMOOTH_JACOBI 0    /* damped point Jacobi                     */
#define CV_SMOOTH_RB     1    /* red-black: damped half-sweeps over even, then odd (i+j+k) cells */
#define CV_SMOOTH_YLINE  2    /* y-line (wall-normal) Jacobi, Thomas solves */

typedef struct cv_ctx cv_ctx;

/* Grid and run parameters: grid.dat sizes, bcs.dat, the MG options of
 * MG_Initial (init.c:296-321, 350-358) and FormInitialize (init.c:52-62). */
typedef struct {
  int IM, JM, KM;          /* grid nodes per axis */
  int bctype[6];           /* i-, i+, j-, j+, k-, k+ ; supported: 1 wall, 7 periodic (i, k),
                              3 slip and 6 far field (every face but j-: the reference's
                              FormMetrics leaves that face's eta at 0), 5 inlet (k- only),
                              4 outlet (k+ only) */
  int mglevels;            /* -mg_level */
  int isc, jsc, ksc;       /* -mg_{i,j,k}_semi */
  double ren, dt, st;      /* -ren, -dt, Strouhal number */
  int coarsen_auto;        /* 0: every level coarsens the axes the semi flags allow
                              (MGDACreate); 1: a level coarsens, among those axes,
                              only the ones with the most cells (anisotropic grids) */
  int inletprofile;        /* -inlet (main.c:23, 887): the analytic inflow profile of a
                              bctype-5 k- face (bcs.c:158-196), one of 1, -1, 2, 4, 7,
                              10, 11; read only when bctype[4] == 5 */
  double inlet_center[3];  /* -x_c -y_c -z_c (main.c:24, 924-926): subtracted from the
                              inlet face centre before the profile is evaluated */
} cv_grid_desc;

/* Poisson solver options (control.dat~:69-87 prefix ps_) */
typedef struct {
  int precision;           /* CV_FP64 / CV_MIXED / CV_MIXED16 */
  int restart;             /* FGMRES restart (-ps_ksp_gmres_restart, default 30) */
  int max_it;              /* -ps_ksp_max_it */
  double rtol, atol;       /* -ps_ksp_rtol, -ps_ksp_atol: stop when ||b-Ax|| <= max(rtol*||b||, atol) */
  int pre_sweeps, post_sweeps, coarse_sweeps;
  int smoother;            /* CV_SMOOTH_* */
  double omega;            /* smoother damping */
} cv_solver_opts;

/* per-step record (the reference's diagnostics: "Summation RHS"
 * poisson.c:1567, "IBMFlux REV" poisson.c:3806/3933, "Maxdiv" main.c:575,
 * CPU_TIME t_mom t_p t_pr solvers.c:1130-1142) */
typedef struct {
  int ksp_its;
  int converged;
  double ksp_rnorm, ksp_bnorm;
  double rhs_sum;
  double maxdiv;
  long maxdiv_idx;
  double ibm_flux0, ibm_area0, ibm_flux1, ibm_area1;
  double t_mom, t_p, t_pr;  /* seconds, device-event timed */
  int imp_its;              /* ImpRK: pseudo-time iterations taken ("!!norm of dU", implicitsolver.c:7190) */
  double imp_lambda;        /* ImpRK: final pseudo-time step factor */
  double imp_normdU, imp_reldU;  /* ImpRK: last ||dU||_inf and its ratio to the first */
  double flux_in, flux_out; /* FluxInSum of InflowFlux and the corrected outlet flux
                               FluxOutSum of FormBCS ("FluxOutRK", implicitsolver.c:82);
                               0 without an inlet / outlet */
} cv_step_stats;

/* momentum scheme of the fractional step (-imp, main.c; Flow_Solver
 * solvers.c:1022-1040) */
#define CV_MOM_RK    0   /* explicit 4-stage RungeKutta, implicitsolver.c:32-119        */
#define CV_MOM_IMPRK 1   /* dual time stepping ImpRK, implicitsolver.c:7094-7256 with
                            CalcRHS 6449-6741 (BDF2 after the first step of a run)      */
typedef struct {
  int scheme;              /* CV_MOM_* */
  int imp_max_it;          /* -imp_MAX_IT (default 50, main.c:21) */
  double imp_atol, imp_rtol;  /* -imp_atol 1e-7, -imp_rtol 1e-4 (main.c:26) */
  double cfl;              /* initial pseudo-time factor lambda (-cfl, default 0.2, init.c:54) */
} cv_momentum_opts;

/* field identifiers for cv_get_field / cv_put_field */
enum {
  CV_F_UCONT = 0, CV_F_UCAT, CV_F_P, CV_F_PHI, CV_F_RHS, CV_F_B, CV_F_NVERT,
  CV_F_AJ, CV_F_CSI, CV_F_ETA, CV_F_ZET,
  CV_F_ICSI, CV_F_IETA, CV_F_IZET, CV_F_IAJ,
  CV_F_JCSI, CV_F_JETA, CV_F_JZET, CV_F_JAJ,
  CV_F_KCSI, CV_F_KETA, CV_F_KZET, CV_F_KAJ,
  CV_F_UCONT_O, CV_F_NFIELDS
};

/* ---- life cycle: MG_Initial (init.c:296-1060) / MG_Finalize (init.c:1063) ---- */
int  cv_ctx_create(const cv_grid_desc *desc, int device, cv_ctx **out);
int  cv_ctx_destroy(cv_ctx *ctx);
int  cv_set_solver(cv_ctx *ctx, const cv_solver_opts *opts);
void cv_default_solver(cv_solver_opts *opts);
int  cv_set_momentum(cv_ctx *ctx, const cv_momentum_opts *opts);
void cv_default_momentum(cv_momentum_opts *opts);
int  cv_last_error(char *buf, int len);
long cv_device_bytes(const cv_ctx *ctx);

/* Node coordinates, IM*JM*KM*3 doubles (x,y,z per node, node (i,j,k) at
 * i + IM*(j + JM*k)), the content of grid.dat (init.c:689-720); builds
 * FormMetrics + Update_Metrics_PBC on every level (metrics.c:4-1229,
 * implicitsolver.c:2174-2573) on the device in FP64. */
int cv_set_geometry(cv_ctx *ctx, const double *xyz_nodes);

/* Fine-level nvert (mx*my*mz), the output of the immersed-body search
 * (ibm.c:1074-1079); coarse levels follow MyNvertRestriction
 * (poisson.c:1662-1773). */
int cv_set_nvert(cv_ctx *ctx, const double *nvert);

/* Immersed-boundary nodes: the output of the immersed-body search
 * (ibm_search_advanced with InterceptionPoint, ibm.c:1952-2420), one
 * record per IB node (IBMInfo, variables.h:36-72).  Storage indices (i, j,
 * k) as in the DMDA arrays.  With nodes set, the momentum scheme and
 * cv_flow_step run ibm_interpolation_advanced(user, ibm, ibi, 1) where the
 * reference does (implicitsolver.c:47-51, 87-90, 7128-7131, 7170-7173;
 * solvers.c:1095-1101). */
typedef struct {
  int i, j, k;                  /* the IB node (ni, nj, nk) */
  int i1, j1, k1, i2, j2, k2, i3, j3, k3;   /* the three cells of the interception triangle */
  double cr1, cr2, cr3;         /* their interpolation weights */
  double d_s, d_i;              /* distance to the surface and to the interception point */
  double ua[3];                 /* surface velocity at the foot point (sum of vertex u * cs) */
  double dun;                   /* Ua.n - Uaold.n along the node's normal (the dU_n/dt term) */
  double nf[3];                 /* surface element normal (free-slip reconstruction) */
} cv_ib_node;
#define CV_IB_NOSLIP   0        /* noslip (wallfunction.c:12-27) */
#define CV_IB_FREESLIP 1        /* freeslip, the -inv option (wallfunction.c:29-46) */
/* n = 0 removes the nodes; cv_set_nvert removes them too (the face and cell
 * lists follow nvert) */
int cv_set_ib_nodes(cv_ctx *ctx, const cv_ib_node *nodes, int n, int mode);

/* field exchange (Ucont_P_Binary_Input/Output, main.c:75-467, in memory) */
int cv_get_field(cv_ctx *ctx, int field, int level, double *out);
int cv_put_field(cv_ctx *ctx, int field, int level, const double *in);

/* ---- hot path: one entry point per reference routine ---- */
int cv_contra2cart(cv_ctx *ctx);                         /* Contra2Cart        rhs.c:25-144      */
int cv_inflow_flux(cv_ctx *ctx, double *flux_in);        /* InflowFlux         bcs.c:99-246: inlet Ucont
                                                            and Ubcs of a bctype-5 k- face; FluxInSum out */
int cv_outflow_flux(cv_ctx *ctx, double *flux_out);      /* OutflowFlux        bcs.c:403-469: flux through
                                                            a bctype-4 k+ face (diagnostic) */
int cv_form_bcs(cv_ctx *ctx);                            /* FormBCS            bcs.c:4436-5888   */
int cv_form_function1(cv_ctx *ctx, double *rhs_out);     /* FormFunction1      rhs.c:2567-3351   */
int cv_rungekutta(cv_ctx *ctx);                          /* RungeKutta         implicitsolver.c:32-119 */
int cv_imprk(cv_ctx *ctx, cv_step_stats *st);            /* ImpRK              implicitsolver.c:7094-7256 */
int cv_calcrhs(cv_ctx *ctx, double *rhs_out);            /* CalcRHS            implicitsolver.c:6449-6741 */
int cv_volumeflux_rev(cv_ctx *ctx, double *flux_area4);  /* VolumeFlux_rev     poisson.c:3712-3944 */
int cv_poisson_rhs(cv_ctx *ctx, double *b_out);          /* PoissonRHS         poisson.c:1509-1576 */
int cv_poisson_apply(cv_ctx *ctx, int level, const double *x, double *y); /* MatMult(A) of PoissonLHSNew poisson.c:600-1507 */
int cv_poisson_solve(cv_ctx *ctx, cv_step_stats *st);    /* PoissonSolver_MG   poisson.c:1775-2017
                                                            and the PoissonSolver_Hypre call site
                                                            solvers.c:1053-1056 */
int cv_update_pressure(cv_ctx *ctx);                     /* UpdatePressure     poisson.c:2908-3248 */
int cv_projection(cv_ctx *ctx);                          /* Projection         poisson.c:2281-2906 */
int cv_divergence(cv_ctx *ctx, double *maxdiv, long *idx); /* Divergence       main.c:469-593    */
/* cv_divergence on device-resident fields, allocating no device memory.  It
 * has no counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  ucont:
 * 3*mx*my*mz doubles (Cmpnts interleaved), aj and nvert: mx*my*mz doubles,
 * all device pointers in the DMDA layout; out: a device pointer to 2
 * doubles, the maximum |div| and its first storage index (the result of
 * cv_divergence on the same fields, bit for bit).  Runs on `stream` (a
 * hipStream_t; 0 for the null stream) of GPU `device` and returns without
 * waiting.  Its block partials live in a fixed device array of 1024
 * (value, index) pairs (16 KiB) per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU race.
 * Storage of 2^31 points or more, or below 3 per axis, is refused with
 * CV_ERR_ARG. */
int cv_divergence_noalloc(int device, const double *ucont, const double *aj, const double *nvert,
                          int mx, int my, int mz, void *stream, double *out);
int cv_ibm_interpolation(cv_ctx *ctx);                   /* ibm_interpolation_advanced(user, ibm, ibi, 1)
                                                            ibm.c:3012-3254 (nothing without IB nodes) */
int cv_flow_step(cv_ctx *ctx, cv_step_stats *st);        /* Flow_Solver        solvers.c:944-1156 + main.c:1123-1145;
                                                            momentum by the scheme of cv_set_momentum */

/* operator plug-in points of the solve, exported for tests (MatShell
 * callbacks poisson.c:1863-1875, nullspace callback poisson.c:1950-1954) */
int cv_mg_restrict(cv_ctx *ctx, int coarse_level, const double *xf, double *xc);  /* MyRestriction  poisson.c:129-226 */
int cv_mg_prolong(cv_ctx *ctx, int fine_level, const double *xc, double *xf);     /* MyInterpolation poisson.c:265-370 */
int cv_nullspace_project(cv_ctx *ctx, int level, double *x);                      /* PoissonNullSpaceFunction poisson.c:372-563 */

/* ---- serving: continuous batching of independent cases on one GPU (no
 * reference counterpart; the reference runs one case per process,
 * main.c:1094-1170) ----
 * Requests queue in arrival order; a request is admitted when one of the
 * server's slots is free and its context (cv_ctx_bytes) fits the HBM budget
 * beside the cases in flight.  An admitted case keeps all its fields
 * resident on the device for its nsteps fractional steps, either on its
 * own HIP stream (cv_flow_step; the cases in flight overlap on the GPU) or,
 * in CV_SERVE_BATCHED mode, as a member of a cv_batch.  The
 * server copies the inputs at submit; results (p, Ucont) and per-step
 * wall-clock latencies stay with the server until cv_server_release.
 * cv_server_destroy finishes the cases in flight and drops queued ones. */
#define CV_CASE_QUEUED  0
#define CV_CASE_RUNNING 1
#define CV_CASE_DONE    2
#define CV_CASE_FAILED  3

typedef struct cv_server cv_server;
typedef struct {
  int status, rc, steps_done;
  long bytes;                          /* device bytes of the case's context */
  double t_submit, t_admit, t_done;    /* seconds since cv_server_create */
  cv_step_stats last;                  /* statistics of the last step */
} cv_case_info;
typedef struct {
  long queued, running, done;
  long bytes_in_use, bytes_peak, budget;
  double time;                         /* seconds since cv_server_create */
} cv_server_info;

/* device bytes one context for this grid and FGMRES restart length holds
 * (-1 for an unsupported description) */
long cv_ctx_bytes(const cv_grid_desc *desc, int restart);
int  cv_server_create(int device, long hbm_budget_bytes, int slots, cv_server **out);   /* streams mode */
/* mode CV_SERVE_STREAMS: one worker thread and stream per slot, each case
 * stepped by cv_flow_step; CV_SERVE_BATCHED: one scheduler advances all
 * admitted cases together, one cv_batch per group of equal solver options,
 * rebuilt when cases join or leave (slots bounds the cases in flight) */
#define CV_SERVE_STREAMS 0
#define CV_SERVE_BATCHED 1
int  cv_server_create_mode(int device, long hbm_budget_bytes, int slots, int mode, cv_server **out);
int  cv_server_submit(cv_server *s, const cv_grid_desc *grid, const cv_solver_opts *opts, const double *xyz_nodes,
                      const double *nvert, const double *ucont, int nsteps, long *case_id);
/* the same with the case's IB node records (cv_set_ib_nodes), copied at submit */
int  cv_server_submit_ib(cv_server *s, const cv_grid_desc *grid, const cv_solver_opts *opts, const double *xyz_nodes,
                         const double *nvert, const double *ucont, int nsteps, const cv_ib_node *ib_nodes,
                         int n_ib_nodes, int ib_mode, long *case_id);
int  cv_server_wait(cv_server *s, long case_id, cv_case_info *info);      /* blocks until done or failed */
int  cv_server_step_ms(cv_server *s, long case_id, double *ms, int *its, int max);  /* returns the count */
int  cv_server_fetch(cv_server *s, long case_id, int field /* CV_F_P, CV_F_UCONT */, double *out);
int  cv_server_release(cv_server *s, long case_id);
int  cv_server_status(cv_server *s, cv_server_info *info);
int  cv_server_destroy(cv_server *s);

/* ---- batched ensemble (no reference counterpart: the reference solves one
 * case per run; each case's PoissonSolver_MG, poisson.c:1775-2017, and
 * Flow_Solver step, solvers.c:944-1156, are what a batch runs for many
 * cases at once) ----
 * A batch groups up to 64 contexts of one device that share solver options
 * (y-line smoother; at most 512 rows per y-line in mixed precision, 256 in
 * FP64).  cv_batch_step advances every case by one fractional step: the
 * momentum and projection parts run per case on each case's own stream, the
 * pressure solve runs once for all cases, one kernel launch per multigrid
 * depth and operation covering every case, FGMRES in lockstep with one host
 * synchronisation per iteration, the preconditioner and apply replayed from
 * a hipGraph.  Each case's fields are bit-identical to cv_flow_step on that
 * case alone.  stats has one entry per case (in the order given); the call
 * returns CV_ERR_NOT_CONVERGED when any case missed its tolerance.  The
 * contexts must not be destroyed or re-optioned while the batch exists;
 * new solver options, IB nodes or nvert make the next batch call return
 * CV_ERR_STATE. */
typedef struct cv_batch cv_batch;
int cv_batch_create(cv_ctx *const *cases, int ncases, cv_batch **out);
int cv_batch_step(cv_batch *b, cv_step_stats *stats);
int cv_batch_poisson_solve(cv_batch *b, cv_step_stats *stats);   /* VolumeFlux_rev + PoissonRHS + solve */
int cv_batch_destroy(cv_batch *b);

/* ---- instrumentation (no reference counterpart; the reference's only
 * timing hook is the per-phase CPU_TIME log, solvers.c:1130-1142) ----
 * Device-event timing of every launch of one kernel class on one level,
 * on the context's stream, while the context runs normally.  level -1 =
 * finest; precision -1 = either, 0 = FP64, 1 = FP32.  cv_prof_read returns
 * the launch count, their summed device time in ms and the summed
 * algorithmic bytes (compulsory HBM traffic of the launch's inputs and
 * outputs); reset clears the counters. */
#define CV_PROF_NONE   0
#define CV_PROF_RESID  1   /* 19-point operator residual r = b - A x       */
#define CV_PROF_APPLY  2   /* 19-point operator apply y = A x (Krylov)     */
#define CV_PROF_LINE   3   /* y-line solve of the line smoother            */
#define CV_PROF_RCT    4   /* momentum face-flux pass (Visc - Conv, Rc -> rct) */
#define CV_PROF_FF1    5   /* momentum face RHS (rct, p -> Rhs)            */
#define CV_PROF_ORTH   6   /* Gram-Schmidt dot products                    */
int cv_prof_select(cv_ctx *ctx, int kernel, int level, int precision);
int cv_prof_read(cv_ctx *ctx, long *launches, double *total_ms, double *bytes, int reset);

#ifdef __cplusplus
}
#endif
#endif
