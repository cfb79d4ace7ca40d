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
  double *red_dev;            /* device result slots (64)                 */
  double *h_red;              /* pinned host result slots (64)            */
};

#define PFHIP_RED_MAX_BLOCKS 1024
#define PFHIP_MAX_MULTI 64

void pfhip_set_error(const char *fmt, ...);
void pfhip_prof_mark(int kernel_id, int end, void *stream);

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
/* MGS step fused: w -= hprev * vprev (if vprev), then returns vnext . w */
int pfhip_vec_axpy_dot(pfhip_problem *pb, double hprev, const double *vprev,
                       double *w, const double *vnext, double *dot,
                       void *stream);

/* difference-quotient J.v in one residual launch (no overland problems):
 * z = sinv F(u + sigma v) - sinv fval */
int pfhip_residual_jv(pfhip_problem *pb, const double *u, const double *v, double sigma,
                      const double *fval, double sinv, double *z, const double *p_old,
                      const double *s_old, double dt, void *stream);

/* RichardsJacobianEval into double (fp32 = 0) or float (fp32 = 1)
 * coefficient arrays J[7] (C, W, E, S, N, L, U; symm: C, E, N, U) */
int pfhip_jacobian_into(pfhip_problem *pb, const double *p, double dt, int symm,
                        void *const *J, int fp32, void *stream);
/* multigrid PC setup from the pressure: symmetric Jacobian straight into the
 * finest level, then the coarse operators */
int pfhip_mg_setup_jacobian(pfhip_mg *mg, const double *p, double dt, void *stream);

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

#ifdef __cplusplus
}
#endif
#endif
