/*
 * pfhip_internal.h -- internal declarations shared by the HIP kernels
 * (*.hip) and the C host driver (kinsol_host.c) of libpfhip.
 */
#ifndef PFHIP_INTERNAL_H
#define PFHIP_INTERNAL_H

#include "../../include/pfhip.h"

#ifdef __cplusplus
extern "C" {
#endif

struct pfhip_problem {
  pfhip_problem_desc d;
  int sy, sz;                 /* strides of the ghosted layout            */
  long ng;                    /* (nx+2)(ny+2)(nz+2)                       */
  long n2;                    /* (nx+2)(ny+2)                             */
  long ncell;                 /* nx*ny*nz                                 */
  int overland;               /* top face is OverlandKinematic            */
  double *bcv[6];             /* device per-face BC values                */
  double *h_bcv[6];           /* pinned host staging                      */
  double *krho, *dkrho;       /* device scratch, ghosted                  */
  double *q2[4];              /* device surface scratch                   */
  double *red_partial;        /* reduction partials (RED_MAX_BLOCKS*8)    */
  double *red_dev;            /* device result slots (PFHIP_RED_SLOTS)    */
  double *h_red;              /* pinned host result slots (64)            */
  double *krtab;              /* device kr lookup table, or NULL          */
  double *stat_partial;       /* KINStop partials of the line-search
                                 residual: 2 x stat_nb                    */
  long stat_nb;               /* residual tiles + boundary-shell blocks   */
};

#define PFHIP_RED_MAX_BLOCKS 1024
#define PFHIP_MAX_MULTI 64
/* device result slots (red_dev) of the fused GMRES steps */
#define PFHIP_SLOT_H2 31          /* CGS2 second-pass coefficients 31..61 */
#define PFHIP_SLOT_N2 62          /* CGS2 ||w||^2                         */
#define PFHIP_SLOT_DQ 64          /* DQ sums of the J.v quotient (3)      */
#define PFHIP_RED_SLOTS 72        /* size of red_dev                      */

void pfhip_set_error(const char *fmt, ...);
void pfhip_prof_mark(int kernel_id, int end, void *stream);
/* nonzero while the per-launch timer of kernel_id is on */
int pfhip_prof_enabled(int kernel_id);

/* ---- extra vector ops used by the nonlinear driver ------------------ */
int pfhip_vec_scale(pfhip_problem *pb, double c, const double *x, double *z,
                    void *stream);
/* max_i |s_i| / (1 + |u_i|)   (KINScSteplength with uscale = 1) */
int pfhip_vec_scsteplength(pfhip_problem *pb, const double *u,
                           const double *s, double *result, void *stream);
/* dot products (u.v, v.v) and ||v||_1 in one pass (KINSpgmrAtimesDQ) */
int pfhip_vec_dq_sums(pfhip_problem *pb, const double *u, const double *v,
                      double *sutsv, double *vtv, double *l1, void *stream);
/* h[j] = V[j] . w for j < m, one pass over w (block Gram-Schmidt) */
int pfhip_vec_multi_dot(pfhip_problem *pb, const double *const *V, int m,
                        const double *w, double *h, void *stream);
/* w -= sum_j h[j] V[j]; optionally returns ||w||^2 after the update */
int pfhip_vec_multi_axpy(pfhip_problem *pb, const double *const *V, int m,
                         const double *h, double *w, double *norm2_out,
                         void *stream);
/* x += sum_j y[j] V[j] */
int pfhip_vec_combine(pfhip_problem *pb, const double *const *V, int m,
                      const double *y, double *x, void *stream);
/* p.p and the scaled step length max |p| / (1 + |u|) in one pass (the
 * l2norm and KINScSteplength of KINLineSearch's start) */
int pfhip_vec_dot_scsteplength(pfhip_problem *pb, const double *u, const double *p,
                               double *pp, double *rlength, void *stream);
/* x = sum y[k] V[k] (zero: x not read) or x +=, with x.x and
 * max |x| / (1 + |u|) from the same pass */
int pfhip_vec_combine_step(pfhip_problem *pb, const double *const *V, int m,
                           const double *y, double *x, const double *u, int zero,
                           double *xx, double *steplength, void *stream);
/* KINSOL's start: max |f|, f.f and u.u in one pass */
int pfhip_vec_kinsol_init_norms(pfhip_problem *pb, const double *u, const double *f,
                                double *fmax, double *ff, double *uu, void *stream);
/* MGS step fused: w -= hprev * vprev (if vprev), then returns vnext . w */
int pfhip_vec_axpy_dot(pfhip_problem *pb, double hprev, const double *vprev,
                       double *w, const double *vnext, double *dot,
                       void *stream);

/* difference-quotient J.v in one residual launch (no overland problems):
 * z = sinv F(u + sigma v) - sinv fval */
int pfhip_residual_jv(pfhip_problem *pb, const double *u, const double *v, double sigma,
                      const double *fval, double sinv, double *z, const double *p_old,
                      const double *s_old, double dt, void *stream);

/* the same with sigma and 1/sigma formed on the device from the DQ sums
 * (u.v, v.v, |v|_1) with sqrt(relfunc) sqrf: no host synchronisation */
int pfhip_residual_jv_dq(pfhip_problem *pb, const double *u, const double *v, double sqrf,
                         const double *fval, double *z, const double *p_old,
                         const double *s_old, double dt, void *stream);
/* pfhip_residual_step plus F.F and KINStop's max |F| and scaled step
 * length max |unew - u| / (1 + |unew|) of the same point, with one host
 * synchronisation (no overland problems) */
int pfhip_residual_step_stats(pfhip_problem *pb, const double *u, const double *v,
                              double sigma, double *unew, double *F, const double *p_old,
                              const double *s_old, double dt, double *ff, double *fmax,
                              double *steplength, void *stream);
/* CGS2 of w against V[0..k-1] (k <= PFHIP_SLOT_H2) with one host
 * synchronisation: h1, h2 and ||w||^2 after the second pass */
int pfhip_vec_cgs2(pfhip_problem *pb, const double *const *V, int k, double *w, double *h1,
                   double *h2, double *norm2, void *stream);
/* the same with basis vector q read as sc[q] * V[q] (sc null: as stored) */
int pfhip_vec_cgs2_scaled(pfhip_problem *pb, const double *const *V, const double *sc, int k,
                          double *w, double *h1, double *h2, double *norm2, void *stream);

/* RichardsJacobianEval into double (fp32 = 0) or float (fp32 = 1)
 * coefficient arrays J[7] (C, W, E, S, N, L, U; symm: C, E, N, U) */
int pfhip_jacobian_into(pfhip_problem *pb, const double *p, double dt, int symm,
                        void *const *J, int fp32, void *stream);
/* multigrid PC setup from the pressure: symmetric Jacobian straight into the
 * finest level, then the coarse operators */
int pfhip_mg_setup_jacobian(pfhip_mg *mg, const double *p, double dt, void *stream);
/* x = M^{-1} (bscale * b), the factor applied where the V-cycle reads b */
int pfhip_mg_apply_scaled(pfhip_mg *mg, const double *b, double bscale, double *x,
                          void *stream);
/* the same, and the J.v sums u.x, x.x, |x|_1 into the problem's DQ slots */
int pfhip_mg_apply_dq(pfhip_mg *mg, const double *b, double bscale, double *x,
                      const double *u, void *stream);
/* the DQ slots' three sums, to the host */
int pfhip_dq_slots(pfhip_problem *pb, double *out, void *stream);
/* difference-quotient J.v with the sums already in the DQ slots */
int pfhip_residual_jv_ready(pfhip_problem *pb, const double *u, const double *v, double sqrf,
                            const double *fval, double *z, const double *p_old,
                            const double *s_old, double dt, void *stream);

/* line-search trial point in one residual launch (no overland problems):
 * unew = uu + rl pp, F = F(unew) */
int pfhip_residual_step(pfhip_problem *pb, const double *uu, const double *pp, double rl,
                        double *unew, double *F, const double *p_old, const double *s_old,
                        double dt, void *stream);
/* KINStop norms in one pass: max|fval| and max|unew-uu|/(1+|unew|) */
int pfhip_kin_stop_norms(pfhip_problem *pb, const double *fval, const double *unew,
                         const double *uu, double *fmax, double *steplength, void *stream);

/* ---- device scratch helpers ---------------------------------------- */
int pfhip_dev_alloc(double **p, long n);
int pfhip_zero(double *p, long n, void *stream);

/* BCPressure values of time `time` into the device tables; sync = 0 leaves
 * the host-to-device copies in flight (the caller guarantees a stream
 * synchronisation before the next call on the same problem) */
int pfhip_bc_pressure_async(pfhip_problem *pb, double time, int sync, void *stream);

/* ---- device operations of the nonlinear driver ------------------------
 * The driver (kinsol_host.c) reaches the device only through this table:
 * the direct table calls the entry points above; the ensemble's table
 * turns each call into a request of a batched launch. */
typedef struct pfhip_ops {
  int (*vec_scale)(pfhip_problem *, double, const double *, double *, void *);
  int (*vec_linear_sum)(pfhip_problem *, double, const double *, double, const double *,
                        double *, void *);
  int (*vec_const)(pfhip_problem *, double, double *, void *);
  int (*vec_dot)(pfhip_problem *, const double *, const double *, double *, void *);
  int (*vec_max_norm)(pfhip_problem *, const double *, double *, void *);
  int (*vec_scsteplength)(pfhip_problem *, const double *, const double *, double *, void *);
  int (*vec_dq_sums)(pfhip_problem *, const double *, const double *, double *, double *,
                     double *, void *);
  int (*vec_multi_dot)(pfhip_problem *, const double *const *, int, const double *,
                       double *, void *);
  int (*vec_multi_axpy)(pfhip_problem *, const double *const *, int, const double *,
                        double *, double *, void *);
  int (*vec_combine)(pfhip_problem *, const double *const *, int, const double *, double *,
                     void *);
  int (*kin_stop_norms)(pfhip_problem *, const double *, const double *, const double *,
                        double *, double *, void *);
  int (*nl_function_eval)(pfhip_problem *, const double *, double *, const double *,
                          const double *, double, void *);
  int (*residual_jv)(pfhip_problem *, const double *, const double *, double, const double *,
                     double, double *, const double *, const double *, double, void *);
  int (*residual_step)(pfhip_problem *, const double *, const double *, double, double *,
                       double *, const double *, const double *, double, void *);
  int (*mg_setup_jacobian)(pfhip_mg *, const double *, double, void *);
  int (*mg_apply)(pfhip_mg *, const double *, double *, void *);
  int (*bc_pressure)(pfhip_problem *, double, void *);
  int (*saturation)(pfhip_problem *, const double *, double *, int, void *);
  int (*synchronize)(void);
  /* fused forms without host synchronisation; NULL: the driver composes
   * them from the operations above */
  int (*residual_jv_dq)(pfhip_problem *, const double *, const double *, double,
                        const double *, double *, const double *, const double *, double,
                        void *);
  int (*cgs2)(pfhip_problem *, const double *const *, int, double *, double *, double *,
              double *, void *);
  int (*vec_dot_scsteplength)(pfhip_problem *, const double *, const double *, double *,
                              double *, void *);
  int (*residual_step_stats)(pfhip_problem *, const double *, const double *, double,
                             double *, double *, const double *, const double *, double,
                             double *, double *, double *, void *);
  int (*kinsol_init_norms)(pfhip_problem *, const double *, const double *, double *,
                           double *, double *, void *);
  int (*vec_combine_step)(pfhip_problem *, const double *const *, int, const double *,
                          double *, const double *, int, double *, double *, void *);
  int (*cgs2_scaled)(pfhip_problem *, const double *const *, const double *, int, double *,
                     double *, double *, double *, void *);
  int (*mg_apply_scaled)(pfhip_mg *, const double *, double, double *, void *);
  int (*mg_apply_dq)(pfhip_mg *, const double *, double, double *, const double *, void *);
  int (*residual_jv_ready)(pfhip_problem *, const double *, const double *, double,
                           const double *, double *, const double *, const double *, double,
                           void *);
} pfhip_ops;

extern const pfhip_ops pfhip_direct_ops;
void pfhip_solver_set_ops(pfhip_solver *s, const pfhip_ops *ops);
pfhip_problem *pfhip_solver_problem(pfhip_solver *s);
pfhip_mg *pfhip_solver_mg(pfhip_solver *s);
const pfhip_solver_opts *pfhip_solver_options(pfhip_solver *s);

/* ---- ensemble (many independent problems per launch) ------------------ */
enum {
  PFHIP_EOP_NONE = 0,
  /* vector operations: one workgroup per member, one launch for all */
  PFHIP_EOP_LIN,        /* z = [z] + sum_q c_q x_q (optionally ||z||^2) */
  PFHIP_EOP_DOT,        /* x . y                                      */
  PFHIP_EOP_MAXABS,     /* max |x|                                    */
  PFHIP_EOP_SCSTEP,     /* max |y| / (1 + |x|)                         */
  PFHIP_EOP_DQ,         /* x . y, y . y, |y|_1                         */
  PFHIP_EOP_MDOT,       /* V_q . y, q < m                              */
  PFHIP_EOP_STOP,       /* max |x|, max |y - u| / (1 + |y|)            */
  /* residual forms: mode 0 F(x), 1 J.v quotient, 2 line-search point */
  PFHIP_EOP_RESIDUAL,
  PFHIP_EOP_MG_SETUP,   /* symmetric Jacobian into level 0 + coarse ops */
  PFHIP_EOP_MG_APPLY,   /* one V-cycle                                 */
  PFHIP_EOP_BC,         /* BCPressure values at time dt                */
  PFHIP_EOP_SAT,        /* Saturation of x into z (mode: CALCFCN/DER)  */
  PFHIP_EOP_COUNT
};

#define PFHIP_ENS_MAXQ 33       /* terms of one EOP_LIN request          */
#define PFHIP_ENS_NRES 32       /* result slots per request              */

typedef struct pfhip_ens_req {
  int op;
  int mode;                     /* residual form                         */
  int base;                     /* EOP_LIN: z itself is the first term   */
  int norm;                     /* EOP_LIN: also ||z||^2                 */
  int nq;                       /* EOP_LIN terms / EOP_MDOT vectors      */
  pfhip_problem *pb;
  pfhip_mg *mg;
  double cst, sig, fsc, dt;
  double c[PFHIP_ENS_MAXQ];
  const double *xq[PFHIP_ENS_MAXQ];
  const double *x, *y, *u;
  double *z, *z2;
  double *scratch[4];           /* member scratch (ghosted, ng doubles)  */
  double *out[PFHIP_ENS_NRES];  /* host destinations of the results      */
  int nout;
  int slot;                     /* result slot within its round          */
  int member;                   /* index of the requesting member        */
  int rc;
} pfhip_ens_req;

/* staging ring for the per-launch member tables: pinned host memory copied
 * asynchronously to its device twin; off returns to 0 after a stream
 * synchronisation */
typedef struct pfhip_ens_stage {
  char *host, *dev;
  long off, cap;
} pfhip_ens_stage;

void *pfhip_ens_upload(pfhip_ens_stage *st, const void *src, long bytes, void *stream);
/* in-place form: reserve returns the host address to fill, commit copies it
 * and returns the device address */
void *pfhip_ens_reserve(pfhip_ens_stage *st, long bytes, void *stream);
void *pfhip_ens_commit(pfhip_ens_stage *st, long bytes, void *stream);
/* batched launches; results of request r land in res[r * PFHIP_ENS_NRES + v] */
int pfhip_ens_exec_vec(pfhip_ens_req *const *rq, int n, double *res, double *part,
                       pfhip_ens_stage *st, void *stream);
/* blocks of the flat vector grid for a member of ncell cells; part needs
 * PFHIP_ENS_NRES doubles per block */
long pfhip_ens_vec_blocks(long ncell);
int pfhip_ens_exec_residual(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                            void *stream);
int pfhip_ens_exec_jacobian(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                            void *stream);
/* the ensemble's multigrid launch plan: which levels of each member run as
 * grid-level phases (one launch per phase covering every member with that
 * level) and which inside one workgroup per member; built once from every
 * member's hierarchy */
typedef struct pfhip_ens_mgplan pfhip_ens_mgplan;
int pfhip_ens_mg_plan_create(pfhip_mg *const *mgs, int n, pfhip_ens_mgplan **out);
int pfhip_ens_mg_plan_destroy(pfhip_ens_mgplan *p);
long pfhip_ens_mg_plan_launches(const pfhip_ens_mgplan *p);
int pfhip_ens_exec_mg_coarse(pfhip_ens_mgplan *p, pfhip_ens_req *const *rq, int n,
                             pfhip_ens_stage *st, void *stream);
int pfhip_ens_exec_mg_apply(pfhip_ens_mgplan *p, pfhip_ens_req *const *rq, int n,
                            pfhip_ens_stage *st, void *stream);
int pfhip_ens_exec_bc(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st, void *stream);
int pfhip_ens_exec_saturation(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                              void *stream);
/* the smoother tile shape chosen for each grid level (its autotuner);
 * returns the number of grid levels written */
int pfhip_mg_pair_shapes(pfhip_mg *mg, int *shape, int max);
/* level-0 matrix arrays (C, E, N, U) and precision of a multigrid */
void pfhip_mg_level0(pfhip_mg *mg, void **A, int *prec);
int pfhip_ens_stage_create(pfhip_ens_stage *st, long cap);
void pfhip_ens_stage_destroy(pfhip_ens_stage *st);
/* n doubles device -> pinned host, then a stream synchronisation */
int pfhip_ens_fetch(double *host, const double *dev, long n, void *stream);
int pfhip_host_alloc(double **p, long n);
void pfhip_host_free(double *p);

#ifdef __cplusplus
}
#endif
#endif
