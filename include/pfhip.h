/*
 * pfhip.h -- C-ABI of the MI355X (gfx950) Richards-equation Newton-Krylov
 * path.  Plain C types only: device pointers, sizes, an opaque stream
 * (hipStream_t passed as void*; NULL = the default stream) and int error
 * codes.  No entry point exits the process; every one returns PFHIP_OK (0)
 * or a negative error code, with a message from pfhip_last_error().
 *
 * Data layout is ParFlow's Subvector layout (pfsimulator/parflow_lib/
 * vector.h:133-137): cell-centred double arrays with one ghost layer, x
 * fastest, (nx+2)*(ny+2)*(nz+2) entries; cell (i,j,k) of the interior is at
 * (i+1) + (j+1)*(nx+2) + (k+1)*(nx+2)*(ny+2).  Surface (2-D) arrays are
 * (nx+2)*(ny+2).  So a single-subgrid ParFlow Vector's SubvectorData()
 * pointer, once resident on the device, can be passed as is.
 *
 * Stencil matrices are 7 separate ghosted coefficient arrays in ParFlow's
 * stencil order C, W, E, S, N, L, U (PL/richards_jacobian_eval.c:126-132).
 * A symmetric matrix (PCMatrixType PFSymmetric) stores only C, E, N, U; its
 * W/S/L couplings are the west/south/lower neighbour's E/N/U
 * (PL/matrix.c:258-290, 470-490).
 */
#ifndef PFHIP_H
#define PFHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PFHIP_OK             0
#define PFHIP_ERR_HIP       -1   /* a HIP runtime call failed            */
#define PFHIP_ERR_ARG       -2   /* invalid argument / shape             */
#define PFHIP_ERR_NOMEM     -3   /* device allocation failed             */
#define PFHIP_ERR_SOLVER    -4   /* nonlinear solve did not converge     */

#define PFHIP_CALCFCN 0
#define PFHIP_CALCDER 1

#define PFHIP_BC_FLUX       0    /* FluxConst                            */
#define PFHIP_BC_DIR_EQUIL  1    /* DirEquilRefPatch                     */
#define PFHIP_BC_OVERLAND   2    /* OverlandKinematic (z-upper face)     */
#define PFHIP_MAX_INTERVALS 8
/* Largest ghosted grid (nx+2)(ny+2)(nz+2) a problem accepts: 2^29 cells, an
 * FP64 array of 4 GiB (the residual kernels address cells by 32-bit byte
 * offsets from each array's base). */
#define PFHIP_MAX_GHOSTED_CELLS 536870912L

/* One box face: type, time cycle (PL/time_cycle_data.c:152-197) and, for
 * DirEquilRefPatch, the reference elevation (PL/problem_bc_pressure.c:143). */
typedef struct {
  int type;
  int n_intervals;
  int repeat;
  int lengths[PFHIP_MAX_INTERVALS];
  double values[PFHIP_MAX_INTERVALS];
  double z_ref;
} pfhip_bc_spec;

/* ProblemData for a box domain (every cell active).  Face order: x-lower,
 * x-upper, y-lower, y-upper, z-lower, z-upper. */
typedef struct {
  int nx, ny, nz;
  double dx, dy, dz;
  double x0, y0, z0;
  double rho, mu, gravity;           /* constant density / viscosity        */
  const double *kx, *ky, *kz;        /* device, ghosted; ghost entries 0    */
  const double *poro, *ss;           /* device, ghosted                     */
  const double *alpha, *nvg;         /* device, ghosted van Genuchten       */
  const double *sres, *ssat;         /* device, ghosted, or NULL for ...    */
  double sres_value, ssat_value;     /* ... these region-constant values    */
  const double *src;                 /* device, ghosted, or NULL            */
  const double *slope_x, *slope_y;   /* device, surface, or NULL            */
  const double *mann;                /* device, surface, or NULL            */
  double ov_eps;                     /* Solver.OverlandKinematic.Epsilon    */
  pfhip_bc_spec bc[6];
  double base_unit, start_time;      /* TimingInfo.BaseUnit / StartTime     */
  /* van Genuchten kr lookup table (Geom.*.RelPerm.NumSamplePoints,
   * .MinPressureHead, .InterpolationMethod; PL/problem_phase_rel_perm.c
   * :116-354, :1894-1915): points 0 for the closed form; else the table of
   * kr and its derivative for the region's alpha and n (which every cell
   * must then share), method 0 Spline, 1 Linear, min_head < 0. */
  int vg_table_points, vg_table_method;
  double vg_table_min_head, vg_table_alpha, vg_table_n;
} pfhip_problem_desc;

typedef struct pfhip_problem pfhip_problem;
typedef struct pfhip_mg pfhip_mg;
typedef struct pfhip_solver pfhip_solver;

/* ---- library / memory ------------------------------------------------- */
const char *pfhip_last_error(void);
int  pfhip_device_count(int *count);
int  pfhip_malloc(void **ptr, long bytes);
int  pfhip_free(void *ptr);
int  pfhip_memcpy_h2d(void *dst, const void *src, long bytes);
int  pfhip_memcpy_d2h(void *dst, const void *src, long bytes);
int  pfhip_memset(void *dst, int value, long bytes);
int  pfhip_synchronize(void);

/* ---- problem ---------------------------------------------------------- */
/* Replaces the ProblemData handed to the Richards modules
 * (PL/problem.c:420-460).  The problem keeps the device pointers of desc
 * (it does not copy the fields) and owns its scratch and BC arrays. */
int  pfhip_problem_create(const pfhip_problem_desc *desc, pfhip_problem **out);
int  pfhip_problem_destroy(pfhip_problem *pb);
/* BCPressure (PL/problem_bc_pressure.c:60): evaluates the six faces' values
 * at `time` on the host and uploads them. */
int  pfhip_bc_pressure(pfhip_problem *pb, double time, void *stream);

/* ---- physics modules -------------------------------------------------- */
/* SaturationInvoke (PL/parflow_proto.h:1033), van Genuchten curve
 * (PL/problem_saturation.c:221-416); fcn = PFHIP_CALCFCN or _CALCDER. */
int  pfhip_saturation(pfhip_problem *pb, const double *p, double *out,
                      int fcn, void *stream);
/* PhaseRelPermInvoke (PL/parflow_proto.h:976), van Genuchten
 * (PL/problem_phase_rel_perm.c:567-1000). */
int  pfhip_phase_rel_perm(pfhip_problem *pb, const double *p, double *out,
                          int fcn, void *stream);
/* NlFunctionEvalInvoke (PL/parflow_proto.h:498; PL/nl_function_eval.c:167).
 * Uses the BC values set by the last pfhip_bc_pressure call. */
int  pfhip_nl_function_eval(pfhip_problem *pb, const double *p, double *F,
                            const double *p_old, const double *s_old,
                            double dt, void *stream);
/* The face velocities NlFunctionEval writes into x/y/z_velocity
 * (PL/nl_function_eval.c:846-848 interior, :1230 DirichletBC, :1432
 * FluxBC, :2023 OverlandKinematicBC; vectors created at
 * PL/solver_richards.c:1034-1044, written out as velx/vely/velz PFBs at
 * :1671-1700),
 * evaluated at pressure p with the BC values of the last
 * pfhip_bc_pressure call.  Face-centred device arrays without ghosts:
 *   vx[i + (nx+1)*(j + ny*k)]    (nx+1)*ny*nz
 *   vy[i + nx*(j + (ny+1)*k)]    nx*(ny+1)*nz
 *   vz[i + nx*(j + ny*k)]        nx*ny*(nz+1)
 * p_old and dt enter only an OverlandKinematic top (q_overlnd / dt);
 * they may be NULL / 0 otherwise.  Overwrites the problem's internal
 * kr*rho buffer. */
int  pfhip_face_velocity(pfhip_problem *pb, const double *p,
                         const double *p_old, double dt, double *vx,
                         double *vy, double *vz, void *stream);
/* RichardsJacobianEvalInvoke (PL/parflow_proto.h:1135;
 * PL/richards_jacobian_eval.c:217).  J[7] device arrays; with symm_part
 * only J[0], J[2], J[4], J[6] are written (J[1], J[3], J[5] may be NULL). */
int  pfhip_richards_jacobian_eval(pfhip_problem *pb, const double *p,
                                  double dt, int symm_part, double *const *J,
                                  void *stream);
/* Matvec (PL/parflow_proto.h:383; PL/matvec.c:40): y = A x. */
int  pfhip_matvec(pfhip_problem *pb, const double *const *A, int symm,
                  const double *x, double *y, void *stream);
/* OverlandFlowEvalKinInvoke (PL/parflow_proto.h:900;
 * PL/overlandflow_eval_Kin.c:60).  CALCFCN: out0 = qx, out1 = qy (face
 * fluxes, surface arrays).  CALCDER: out0..3 = KE, KW, KN, KS derivatives. */
int  pfhip_overland_flow_eval_kin(pfhip_problem *pb, const double *p,
                                  double *out0, double *out1, double *out2,
                                  double *out3, int fcn, void *stream);

/* ---- N_Vector operations (PL/n_vector.h -> PL/vector_utilities.c) ---- */
/* All act on the interior cells of ghosted vectors of problem pb (ghost
 * entries are neither read nor written); reductions are deterministic
 * (fixed-order block partials) and return their result on the host. */
/* PFVLinearSum (:80): z = a x + b y */
int  pfhip_vec_linear_sum(pfhip_problem *pb, double a, const double *x,
                          double b, const double *y, double *z, void *stream);
/* PFVConstInit (:233): z = c */
int  pfhip_vec_const(pfhip_problem *pb, double c, double *z, void *stream);
/* PFVProd (:280): z = x y */
int  pfhip_vec_prod(pfhip_problem *pb, const double *x, const double *y,
                    double *z, void *stream);
/* PFVDiv (:352): z = x / y */
int  pfhip_vec_div(pfhip_problem *pb, const double *x, const double *y,
                   double *z, void *stream);
/* PFVScale (:423): z = c x */
int  pfhip_vec_scale(pfhip_problem *pb, double c, const double *x, double *z,
                     void *stream);
/* PFVAbs (:500): z = |x| */
int  pfhip_vec_abs(pfhip_problem *pb, const double *x, double *z, void *stream);
/* PFVInv (:558): z = 1 / x */
int  pfhip_vec_inv(pfhip_problem *pb, const double *x, double *z, void *stream);
/* PFVAddConst (:617): z = x + b */
int  pfhip_vec_add_const(pfhip_problem *pb, const double *x, double b,
                         double *z, void *stream);
/* PFVDotProd (:677): sum x y */
int  pfhip_vec_dot(pfhip_problem *pb, const double *x, const double *y,
                   double *result, void *stream);
/* PFVMaxNorm (:748): max |x| */
int  pfhip_vec_max_norm(pfhip_problem *pb, const double *x, double *result,
                        void *stream);
/* PFVWrmsNorm (:805): sqrt(sum (x w)^2 / N), N the interior cell count */
int  pfhip_vec_wrms_norm(pfhip_problem *pb, const double *x, const double *w,
                         double *result, void *stream);
/* PFVWL2Norm (:877): sqrt(sum (x w)^2) */
int  pfhip_vec_wl2_norm(pfhip_problem *pb, const double *x, const double *w,
                        double *result, void *stream);
/* PFVL1Norm (:949): sum |x| */
int  pfhip_vec_l1_norm(pfhip_problem *pb, const double *x, double *result,
                       void *stream);
/* PFVMin (:1005): min x */
int  pfhip_vec_min(pfhip_problem *pb, const double *x, double *result,
                   void *stream);
/* PFVConstrProdPos (:1145): *result = 0 (FALSE) if some c != 0 has
 * x c <= 0, else 1 */
int  pfhip_vec_constr_prod_pos(pfhip_problem *pb, const double *c,
                               const double *x, int *result, void *stream);
/* PFVCompare (:1228): z = (|x| >= c) ? 1 : 0 */
int  pfhip_vec_compare(pfhip_problem *pb, double c, const double *x, double *z,
                       void *stream);
/* PFVInvTest (:1288): z = 1 / x where x != 0 (unchanged where x == 0);
 * *result = 0 (FALSE) if some x == 0, else 1 */
int  pfhip_vec_inv_test(pfhip_problem *pb, const double *x, double *z,
                        int *result, void *stream);

/* ---- water balance ------------------------------------------------------ */
/* Water-balance terms of one step, as volumes over the step, into host
 * out[PFHIP_WB_TERMS]: [0] inflow through the domain faces other than an
 * OverlandKinematic top (-dt sum dir h v, v the BC face velocity of
 * PL/nl_function_eval.c:1230/:1432); [1] inflow of the overland top's flux
 * BC (rain, -dt h value); [2] surface runoff leaving through the domain
 * edges (dt times the kinematic edge fluxes, PL/overlandflow_eval_Kin.c);
 * [3] sources, sum src vol dt (the increment EvapTransSum adds,
 * PL/evaptranssum.c:31).  The budget the reference's water-balance tests
 * close (test/python/water_balance_x.py:546-620): for a converged step
 * the change of stored water -- the residual's accumulation terms summed
 * (phi S rho vol + Ss vol p S rho) plus the surface storage dx dy
 * max(p_top, 0) -- equals [0] + [1] - [2] + [3] to the Newton tolerance.
 * pftools' subsurface storage (pfhip_water_balance_storage) drops the
 * compressible term where p < 0, so its change closes the budget only
 * where that term is small.  PL/overlandsum.c:33 (OverlandSum) accumulates
 * only for OverlandFlow patches (type 7), never for OverlandKinematic, so
 * the runoff term here is what the kinematic path has in its place.
 * p_old: the step's initial pressure (overland top only).  Overwrites the
 * problem's internal kr*rho buffer. */
#define PFHIP_WB_TERMS 4
int  pfhip_water_balance_fluxes(pfhip_problem *pb, const double *p,
                                const double *p_old, double dt, double *out,
                                void *stream);
/* EvapTransSum (PL/evaptranssum.c:31; called at PL/solver_richards.c:3680):
 * sum += src * dx*dy*dz * dt on every cell of the ghosted device vector
 * sum.  No-op for a problem without a source field. */
int  pfhip_evap_trans_sum(pfhip_problem *pb, double dt, double *sum,
                          void *stream);
/* calculate_subsurface_storage / calculate_surface_storage
 * (pftools/python/parflow/tools/hydrology.py:136-205) summed over the
 * domain: subsurface = sum(phi S dz dx dy + [p >= 0] p S Ss dz dx dy),
 * surface = sum over the top layer of max(p dx dy, 0).  p, s device. */
int  pfhip_water_balance_storage(pfhip_problem *pb, const double *p,
                                 const double *s, double *subsurface,
                                 double *surface, void *stream);
/* No counterpart in the reference.  pfhip_water_balance_storage for
 * callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and the same results
 * bit for bit, with the reduction partials and results kept in a
 * file-scope device array of PFHIP_NOALLOC_SCRATCH_DOUBLES doubles, one
 * copy per GPU, held until the process ends; this call allocates nothing.
 * Its address is looked up per device at pfhip_problem_create (or the
 * first call on a device).  Calls that overlap on one GPU (several
 * streams or host threads) share that copy and race: serialise them.
 * A problem whose reduction needs more scratch is rejected with
 * PFHIP_ERR_ARG (none does today: each sum has at most 1024 block
 * partials, so 2 * 1024 + 2 doubles are used at most). */
#define PFHIP_NOALLOC_SCRATCH_DOUBLES 4096
int  pfhip_water_balance_storage_noalloc(pfhip_problem *pb, const double *p,
                                         const double *s, double *subsurface,
                                         double *surface, void *stream);

/* ---- multigrid preconditioner (PrecondInvoke, PL/parflow_proto.h:418) - */
/* Semicoarsening V-cycle with the level structure, operator-induced
 * transfers and Galerkin-collapsed coarse operators of MGSemi
 * (PL/mg_semi.c:93-760), red/black Gauss-Seidel smoothing
 * (PL/rb_GS_point.c:56).  precision = 64 (FP64) or 32 (FP32 levels). */
int  pfhip_mg_create(pfhip_problem *pb, int max_levels, int pre, int post,
                     int symmetric, int precision, pfhip_mg **out);
/* The same hierarchy for PCMatrixType FullJacobian (PL/kinsol_pc.c:215-240,
 * the non-symmetric hypre matrix of PL/pf_hypre.c:300-350 and 440-520):
 * every level keeps all seven couplings (C, W, E, S, N, L, U).  Coarse
 * operators are P^T A P along the coarsening direction with each side of
 * a row collapsed from its own couplings; on a symmetric matrix they equal
 * pfhip_mg_create's. */
int  pfhip_mg_create_full(pfhip_problem *pb, int max_levels, int pre, int post,
                          int symmetric, int precision, pfhip_mg **out);
int  pfhip_mg_destroy(pfhip_mg *mg);
int  pfhip_mg_num_levels(pfhip_mg *mg, int *levels);
/* PrecondInitInstanceXtraInvoke (PL/parflow_proto.h:419): coarse levels
 * from the PC matrix: A[7] in stencil order C, W, E, S, N, L, U; a
 * symmetric hierarchy reads C, E, N, U, a full one all seven. */
int  pfhip_mg_setup(pfhip_mg *mg, const double *const *Asym, void *stream);
/* One V-cycle, zero initial guess: x = M^{-1} b. */
int  pfhip_mg_apply(pfhip_mg *mg, const double *b, double *x, void *stream);

/* ---- nonlinear solver (NonlinSolverInvoke, PL/parflow_proto.h:301) ---- */
typedef struct {
  double residual_tol, step_tol;     /* Solver.Nonlinear.ResidualTol/StepTol */
  int max_iter;                      /* Solver.Nonlinear.MaxIter             */
  int krylov_dim, max_restarts;      /* Solver.Linear.KrylovDimension/...    */
  int eta_choice;                    /* 0 EtaConstant, 1 Walker1, 2 Walker2  */
  double eta_value, eta_alpha, eta_gamma;
  int use_jacobian;                  /* analytic J.v (KINSolMatVec)          */
  double deriv_eps;                  /* DerivativeEpsilon                    */
  int globalization;                 /* 0 InexactNewton, 1 LineSearch        */
  int precond;                       /* 0 none, 1 multigrid                  */
  int mg_pre, mg_post, mg_symmetric, mg_max_levels;
  int mg_precision;                  /* 64 or 32                             */
  int krylov;                        /* 0 SPGMR + modified Gram-Schmidt (the
                                        reference's), 1 flexible GMRES with
                                        block classical Gram-Schmidt (CGS2)  */
  int pc_matrix;                     /* Solver.Linear.Preconditioner.
                                        PCMatrixType: 0 PFSymmetric,
                                        1 FullJacobian                       */
} pfhip_solver_opts;

typedef struct {
  int ret;                           /* KINSOL return code (KS/kinsol.h:469) */
  int nni, nli, nfe, npe, nps, nbcf, nbktrk, ncfl;
  double fnorm;
} pfhip_kin_stats;

int  pfhip_solver_create(pfhip_problem *pb, const pfhip_solver_opts *opts,
                         pfhip_solver **out);
int  pfhip_solver_destroy(pfhip_solver *s);
/* KinsolNonlinSolver (PL/kinsol_nonlin_solver.c:297): Newton solve of one
 * time step in place on p.  Returns PFHIP_OK when the reference would treat
 * the step as converged (KINSOL_SUCCESS / KINSOL_INITIAL_GUESS_OK),
 * PFHIP_ERR_SOLVER otherwise; stats are filled either way. */
int  pfhip_kinsol_nonlin_solver(pfhip_solver *s, double *p,
                                const double *p_old, const double *s_old,
                                double dt, double time, pfhip_kin_stats *st);

/* AdvanceRichards per-step contract (PL/solver_richards.c:2970-3400,
 * 3661-3669): nsteps steps of size dt from t0 with dt halving on failure.
 * p, s are device arrays (in/out).  newton/gmres (host, may be NULL) get
 * per-step iteration counts. */
int  pfhip_advance_richards(pfhip_solver *s, double *p, double *sat,
                            double t0, double dt, int nsteps,
                            int max_conv_failures, int *newton, int *gmres,
                            int *fevals, double *dt_used, int *steps_done);

/* The surface predictor of AdvanceRichards (PL/solver_richards.c:3236-3362;
 * keys Solver.SurfacePredictor, .PressureValue, .LateralFlows at
 * :6959-6976), applied to the Newton initial guess p before a step's solve:
 * a top cell with p < 0 whose free pore volume is smaller than the step's
 * inflow -- src dx dy dz dt, plus dt times the Darcy inflow through its lower
 * and lateral faces (face velocities vx, vy, vz in the layout of
 * pfhip_face_velocity), plus with lateral_flows the surface term
 * dt ((qx_i - qx_(i-1)) dy + (qy_j - qy_(j-1)) dx) as :3312-3321 add it,
 * from qx, qy (the ke, kn of pfhip_overland_flow_eval_kin; both NULL for
 * none) -- gets p = excess / (dx dy phi), or pressure_value when that is
 * > 0.  sat: S at the step's start.  The problem's source field stands for
 * ParFlow's evap_trans here. */
int  pfhip_surface_predictor(pfhip_problem *pb, double *p, const double *sat,
                             const double *vx, const double *vy,
                             const double *vz, const double *qx,
                             const double *qy, double dt,
                             double pressure_value, int lateral_flows,
                             void *stream);

/* ---- time-step control (SelectTimeStep, the AdvanceRichards loop) ------ */
#define PFHIP_TS_CONSTANT 0
#define PFHIP_TS_GROWTH   1
/* TimeStep.* keys (PL/select_time_step.c:200-245). */
typedef struct {
  int type;                          /* PFHIP_TS_CONSTANT / PFHIP_TS_GROWTH  */
  double value;                      /* Constant: TimeStep.Value             */
  double initial_step, factor;       /* Growth: TimeStep.InitialStep,        */
  double min_step, max_step;         /*   GrowthFactor, MinStep, MaxStep     */
} pfhip_time_step;

/* TimingInfo.* and the loop limits of AdvanceRichards. */
typedef struct {
  double start_time, stop_time;      /* TimingInfo.StartTime / StopTime      */
  double base_unit;                  /* TimingInfo.BaseUnit                  */
  double dump_interval;              /* TimingInfo.DumpInterval: > 0 a time
                                        interval, < 0 every -n iterations    */
  int max_iterations;                /* Solver.MaxIter                       */
  pfhip_time_step ts;
  int surface_predictor;             /* Solver.SurfacePredictor              */
  double surface_predictor_pressure; /*   .PressureValue (default 1e-5)      */
  int surface_lateral_flows;         /*   .LateralFlows                      */
} pfhip_time_control;

/* What persists between steps (instance_xtra of PL/solver_richards.c):
 * zero-initialise, then set t = start_time. */
typedef struct {
  double t;                          /* current time                         */
  double dt;                         /* last step size; 0 before the first   */
  int iteration_number;              /* completed steps                      */
  int dump_index;                    /* next dump time index (starts at 1)   */
} pfhip_time_state;

/* One row of the reference's time log (PL/solver_richards.c:4588-4601). */
typedef struct {
  double time, dt;
  char dt_info;                      /* 'p' policy, 'b' BC cycle, 'f' final  */
  int dumped;                        /* step ends on a dump time             */
  int newton, gmres, fevals;         /* summed over the step's attempts      */
  int conv_failures;
} pfhip_step_log;

/* TimeCycleDataComputeNextTransition (PL/time_cycle_data.c:206-299) over
 * the cycles of bc[0..nbc): the time from `time` to the next interval
 * change of any cycle with more than one interval, or -1 when there is
 * none (or it is closer than TIME_EPSILON).  Host only. */
double pfhip_next_transition(const pfhip_bc_spec *bc, int nbc, double base_unit,
                             double start_time, double stop_time, double time);
/* SelectTimeStep (PL/select_time_step.c:58-156): *dt in/out (the previous
 * step, 0 before the first), then clipped to the next BC-cycle transition
 * (*dt_info 'b') or left as chosen ('p').  No wells in this path, so the
 * well transition never binds.  Host only. */
int  pfhip_select_time_step(const pfhip_time_control *tc, const pfhip_bc_spec *bc,
                            int nbc, double time, double *dt, char *dt_info);
/* The AdvanceRichards time loop (PL/solver_richards.c:2081-2090,
 * 2970-3240, 3380-3410, 4603-4612): per step SelectTimeStep (or halving
 * after a failed solve), the dump-interval and stop-time clips, the Newton
 * solve at the new time.  Runs until stop_time, max_iterations or
 * max_steps steps; state carries t, dt and the counters across calls.
 * log (host, max_steps rows, may be NULL) gets one row per step. */
int  pfhip_advance_richards_ts(pfhip_solver *s, double *p, double *sat,
                               const pfhip_time_control *tc, pfhip_time_state *state,
                               int max_steps, int max_conv_failures,
                               pfhip_step_log *log, int *steps_done);

/* ---- ensembles of independent problems ------------------------------- */
/* No single counterpart in the reference, which solves one problem per
 * run: each member keeps the AdvanceRichards contract of
 * pfhip_advance_richards (PL/solver_richards.c:2970-3400) and its own
 * solver, while the device work of all members is batched -- every
 * vector operation, residual form, PC setup and V-cycle that members have
 * pending is served by one launch, and a member that has finished its
 * steps simply stops contributing (a per-member convergence mask).
 * Members must use the flexible Krylov variant with a preconditioner, the
 * difference-quotient J.v and no overland flow; krylov_dim <= 31. */
typedef struct pfhip_ensemble pfhip_ensemble;
int  pfhip_ensemble_create(int n, pfhip_solver *const *solvers, pfhip_ensemble **out);
int  pfhip_ensemble_destroy(pfhip_ensemble *e);
/* nsteps steps of size dt[m] from t0[m] for every member m, on device
 * arrays p[m], sat[m]; newton, gmres, fevals (host, n * nsteps, may be NULL)
 * get per-member per-step counts, steps_done (n, may be NULL) the steps each
 * member completed.  Returns the first member error (its message in
 * pfhip_last_error), after every member has stopped. */
int  pfhip_ensemble_advance_richards(pfhip_ensemble *e, double *const *p,
                                     double *const *sat, const double *t0,
                                     const double *dt, int nsteps,
                                     int max_conv_failures, int *newton, int *gmres,
                                     int *fevals, int *steps_done);
/* scheduler rounds and batched launches of the last advance */
int  pfhip_ensemble_stats(pfhip_ensemble *e, long *rounds, long *launches);

/* ---- live kernel and phase timing (HIP events on the launching stream) */
/* No counterpart in the reference (its timers are host wall clocks,
 * PL/timing.c:75).  kernel_id: a kernel ... */
#define PFHIP_PROF_RESIDUAL  0    /* the residual kernel of NlFunctionEval  */
#define PFHIP_PROF_CLOSURE   1    /* van Genuchten closure kernel           */
#define PFHIP_PROF_JACOBIAN  2    /* RichardsJacobianEval assembly kernel   */
#define PFHIP_PROF_RBGS      3    /* finest-level red/black smoothing launch
                                     with level-precision operands          */
#define PFHIP_PROF_RESIDUAL_JV 4  /* residual kernel in the J.v quotient form */
#define PFHIP_PROF_RESIDUAL_STEP 5 /* residual kernel at a line-search point */
#define PFHIP_PROF_RBGS_DOUT 6    /* the V-cycle's last finest-level pair,
                                     writing the FP64 result               */
#define PFHIP_PROF_RBGS_DIN  7    /* the V-cycle's first finest-level pair,
                                     reading the FP64 right-hand side      */
/* ... or a phase of the nonlinear solve, after the reference's timers: */
#define PFHIP_PROF_PH_NL_F_EVAL 8 /* residual evaluations, F and line-search
                                     points (NlFunctionEval timer,
                                     PL/nl_function_eval.c:293)            */
#define PFHIP_PROF_PH_JV     9    /* J.v products of GMRES (KINSpgmrAtimes:
                                     difference quotient or Matvec,
                                     PL/matvec.c:89)                        */
#define PFHIP_PROF_PH_PC_SETUP 10 /* PC matrix and multigrid setup
                                     (KinsolPC init, PL/kinsol_pc.c:97)    */
#define PFHIP_PROF_PH_PC_SOLVE 11 /* V-cycle applications (PFMG timer,
                                     PL/pf_pfmg.c:98)                      */
#define PFHIP_PROF_PH_ORTHO  12   /* Gram-Schmidt of the Krylov basis
                                     (SpgmrSolve's ModifiedGS, or CGS2)    */
#define PFHIP_PROF_PH_KINSOL 13   /* whole nonlinear solves (KINSol timer,
                                     PL/kinsol_nonlin_solver.c:368)        */
#define PFHIP_PROF_NKERNELS  14
int  pfhip_prof_enable(int kernel_id, int max_launches);
int  pfhip_prof_reset(int kernel_id);
int  pfhip_prof_read(int kernel_id, double *total_ms, long *launches);
int  pfhip_prof_disable(int kernel_id);

#ifdef __cplusplus
}
#endif
#endif
