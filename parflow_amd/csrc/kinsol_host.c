/*
 * kinsol_host.c -- host driver of the Newton-Krylov solve on device vectors.
 *
 * Restates the embedded KINSOL that ParFlow builds when SUNDIALS is absent:
 *   KINSolInit KS/kinsol.c:363-650, KINSol loop :690-860,
 *   KINForcingTerm :1054-1090, KINInitialStop :1128-1137,
 *   KINInexactNewton :1152-1220, KINLineSearch :1228-1405,
 *   KINLinSolDrv :1424-1470, KINStop :1507-1580,
 *   KINSpgmrSolve / KINSpgmrAtimesDQ KS/kinspgmr.c:344-531,
 *   SpgmrSolve KS/spgmr.c:168-450, ModifiedGS/QRfact/QRsol KS/iterativ.c,
 * with ParFlow's callbacks (PL/kinsol_nonlin_solver.c:144-236,
 * PL/kinsol_pc.c:58-178, PL/richards_jacobian_eval.c:165-211) and the
 * AdvanceRichards step contract (PL/solver_richards.c:2970-3400).
 *
 * Every vector lives on the device; this file only sees scalars.  Two
 * Krylov variants:
 *   krylov = 0  right-preconditioned GMRES with modified Gram-Schmidt, the
 *               reference's algorithm (each MGS step is one fused
 *               axpy+dot pass, same arithmetic order);
 *   krylov = 1  flexible GMRES with block classical Gram-Schmidt applied
 *               twice (CGS2): two multi-dot passes per iteration instead of
 *               l+1 dependent dots, and the preconditioned vectors kept so
 *               an FP32 V-cycle can change between applications.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfhip_internal.h"

struct pfhip_solver {
  pfhip_problem *pb;
  const pfhip_ops *ops;         /* device operations (direct or batched)  */
  pfhip_solver_opts o;
  long ng;
  double *unew, *fval, *pp, *vtemp1, *vtemp2, *xcor, *vtemp;
  double *unew_own;             /* unew's allocation (unew and uu exchange) */
  double **V, **Z;
  double *V0_own;               /* V[0]'s buffer (V[0] may point at fval during a solve) */
  double *vsc;                  /* flexible variant: V[k] holds w_k unnormalised and
                                   vsc[k] its factor (1/norm), applied where it is read */
  double *J[7];
  pfhip_mg *mg;
  double *p_old, *s_old;        /* advance_richards state                 */
  double *vel[3];               /* face velocities for the surface predictor */
  /* host Krylov state */
  double **Hes, *givens, *yg, *hbuf;
  double *Hraw;                 /* flexible variant: the Hessenberg columns
                                   before the QR update, (l_max+1) x l_max   */
  int jp_ok;                    /* flexible variant: J p from the Arnoldi
                                   relation of the last solve is valid      */
  double jp_norm2, jp_fdot;     /* its ||J p||^2 and F.(J p)                */
  int l_max, max_restarts;
  /* KINSOL scalars */
  double *uu;
  const double *cur_p_old, *cur_s_old;
  double dt;
  double fnorm, f1norm, eta, ealpha, egamma, sJpnorm, sfdotJp, stepl;
  double mxnewtstep, sqrt_relfunc, uround, fnormtol, scsteptol, eps, pthrsh;
  int etaflag, callForcingTerm, nni, nnilpre, nfe, nbcf, nbktrk, ncscmx;
  int precondcurrent, mxiter, nli, nps, npe, ncfl, new_uu;
  /* KINStop's max |fval| and scaled step length of the current (unew,
   * fval), when the residual launch that produced them also gathered them */
  int stop_valid;
  double stop_fmax, stop_steplength;
  /* the vector whose difference-quotient sums (with uu) the preconditioner
   * left in the problem's DQ slots, for the J.v that follows it */
  const double *dq_ready;
  /* pp.pp and pp's scaled step length from uu, gathered by the pass that
   * formed pp (the flexible GMRES's solution update) */
  int pp_stats_valid;
  double pp_dot, pp_steplength;
};

#define CHK(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

const pfhip_ops pfhip_direct_ops = {
  pfhip_vec_scale, pfhip_vec_linear_sum, pfhip_vec_const, pfhip_vec_dot,
  pfhip_vec_max_norm, pfhip_vec_scsteplength, pfhip_vec_dq_sums, pfhip_vec_multi_dot,
  pfhip_vec_multi_axpy, pfhip_vec_combine, pfhip_kin_stop_norms, pfhip_nl_function_eval,
  pfhip_residual_jv, pfhip_residual_step, pfhip_mg_setup_jacobian, pfhip_mg_apply,
  pfhip_bc_pressure, pfhip_saturation, pfhip_synchronize, pfhip_residual_jv_dq,
  pfhip_vec_cgs2, pfhip_vec_dot_scsteplength, pfhip_residual_step_stats,
  pfhip_vec_kinsol_init_norms, pfhip_vec_combine_step, pfhip_vec_cgs2_scaled,
  pfhip_mg_apply_scaled, pfhip_mg_apply_dq, pfhip_residual_jv_ready,
};

void pfhip_solver_set_ops(pfhip_solver *s, const pfhip_ops *ops) { s->ops = ops; }
pfhip_problem *pfhip_solver_problem(pfhip_solver *s) { return s->pb; }
pfhip_mg *pfhip_solver_mg(pfhip_solver *s) { return s->mg; }
const pfhip_solver_opts *pfhip_solver_options(pfhip_solver *s) { return &s->o; }

static int alloc_vec(pfhip_solver *s, double **p)
{
  CHK(pfhip_dev_alloc(p, s->ng));
  return pfhip_zero(*p, s->ng, NULL);
}

/* the vectors the preconditioner reads and writes (the Krylov basis, fval,
 * which stands in for the first basis vector, the flexible basis Z, xcor
 * and vtemp) start one element past the allocation,
 * so the value pairs of even-i cells (odd ghosted index) are 16-byte aligned
 * and the multigrid moves them with one 16-byte access per lane */
static int alloc_vec_pc(pfhip_solver *s, double **p)
{
  double *base;
  CHK(pfhip_dev_alloc(&base, s->ng + 1));
  *p = base + 1;
  return pfhip_zero(base, s->ng + 1, NULL);
}

static void free_vec_pc(double *p)
{
  if (p) pfhip_free(p - 1);
}

int pfhip_solver_destroy(pfhip_solver *s)
{
  if (!s) return PFHIP_OK;
  double *v[6] = { s->unew_own, s->pp, s->vtemp1, s->vtemp2, s->p_old, s->s_old };
  for (int i = 0; i < 6; i++) pfhip_free(v[i]);
  free_vec_pc(s->fval);
  free_vec_pc(s->xcor);
  free_vec_pc(s->vtemp);
  if (s->V) for (int i = 0; i <= s->l_max; i++) free_vec_pc(s->V[i]);
  if (s->Z) for (int i = 0; i < s->l_max; i++) free_vec_pc(s->Z[i]);
  for (int q = 0; q < 7; q++) pfhip_free(s->J[q]);
  for (int q = 0; q < 3; q++) pfhip_free(s->vel[q]);
  if (s->mg) pfhip_mg_destroy(s->mg);
  if (s->Hes) for (int i = 0; i <= s->l_max; i++) free(s->Hes[i]);
  free(s->Hes); free(s->givens); free(s->yg); free(s->hbuf); free(s->Hraw); free(s->vsc);
  free(s->V); free(s->Z);
  free(s);
  return PFHIP_OK;
}

int pfhip_solver_create(pfhip_problem *pb, const pfhip_solver_opts *opts,
                        pfhip_solver **out)
{
  if (!pb || !opts || !out) { pfhip_set_error("null argument"); return PFHIP_ERR_ARG; }
  if (opts->krylov_dim < 1 || opts->krylov_dim > PFHIP_MAX_MULTI - 2) {
    pfhip_set_error("krylov_dim must be 1..%d", PFHIP_MAX_MULTI - 2);
    return PFHIP_ERR_ARG;
  }
  pfhip_solver *s = (pfhip_solver *)calloc(1, sizeof(pfhip_solver));
  s->pb = pb;
  s->ops = &pfhip_direct_ops;
  s->o = *opts;
  s->ng = pb->ng;
  s->l_max = opts->krylov_dim;
  s->max_restarts = opts->max_restarts < 0 ? 0 : opts->max_restarts;
  int rc = PFHIP_OK;
  double **vv[6] = { &s->unew, &s->pp, &s->vtemp1, &s->vtemp2, &s->p_old, &s->s_old };
  for (int i = 0; i < 6 && !rc; i++) rc = alloc_vec(s, vv[i]);
  if (!rc) rc = alloc_vec_pc(s, &s->fval);
  if (!rc) rc = alloc_vec_pc(s, &s->xcor);
  if (!rc) rc = alloc_vec_pc(s, &s->vtemp);
  s->unew_own = s->unew;
  s->V = (double **)calloc(s->l_max + 1, sizeof(double *));
  for (int i = 0; i <= s->l_max && !rc; i++) rc = alloc_vec_pc(s, &s->V[i]);
  s->V0_own = s->V[0];
  s->vsc = (double *)calloc(s->l_max + 1, sizeof(double));
  if (opts->krylov == 1) {
    s->Z = (double **)calloc(s->l_max, sizeof(double *));
    for (int i = 0; i < s->l_max && !rc; i++) rc = alloc_vec_pc(s, &s->Z[i]);
  }
  /* the PC matrix is assembled straight into the multigrid's finest level */
  for (int q = 0; q < 7 && !rc; q++)
    if (opts->use_jacobian) rc = alloc_vec(s, &s->J[q]);
  if (!rc && opts->precond)
    rc = (opts->pc_matrix ? pfhip_mg_create_full : pfhip_mg_create)(
        pb, opts->mg_max_levels, opts->mg_pre, opts->mg_post, opts->mg_symmetric,
        opts->mg_precision ? opts->mg_precision : 64, &s->mg);
  s->Hes = (double **)calloc(s->l_max + 1, sizeof(double *));
  for (int i = 0; i <= s->l_max; i++) s->Hes[i] = (double *)calloc(s->l_max, sizeof(double));
  s->givens = (double *)calloc(2 * s->l_max, sizeof(double));
  s->yg = (double *)calloc(s->l_max + 1, sizeof(double));
  s->hbuf = (double *)calloc(2 * (s->l_max + 2), sizeof(double));
  s->Hraw = (double *)calloc((size_t)(s->l_max + 1) * s->l_max, sizeof(double));
  if (!rc) rc = s->ops->synchronize();
  if (rc) { pfhip_solver_destroy(s); return rc; }
  *out = s;
  return PFHIP_OK;
}

/* ---- callbacks -------------------------------------------------------- */
/* phase timers (pfhip_prof_*, PFHIP_PROF_PH_*): an event pair around each
 * call of a solver phase when that timer is enabled */
#define PH_BEGIN(id) pfhip_prof_mark((id), 0, NULL)
#define PH_END(id) pfhip_prof_mark((id), 1, NULL)

static int func_eval(pfhip_solver *s, const double *u, double *f)
{
  s->nfe++;
  return s->ops->nl_function_eval(s->pb, u, f, s->cur_p_old, s->cur_s_old, s->dt, NULL);
}

static int func(pfhip_solver *s, const double *u, double *f)
{
  PH_BEGIN(PFHIP_PROF_PH_NL_F_EVAL);
  int rc = func_eval(s, u, f);
  PH_END(PFHIP_PROF_PH_NL_F_EVAL);
  return rc;
}

static int l2norm(pfhip_solver *s, const double *x, double *r)
{
  double d;
  CHK(s->ops->vec_dot(s->pb, x, x, &d, NULL));
  *r = sqrt(d);
  return PFHIP_OK;
}

/* KINSolInitPC -> KinsolPCInitInstanceXtra: PC matrix + multigrid setup */
static int pc_setup(pfhip_solver *s)
{
  PH_BEGIN(PFHIP_PROF_PH_PC_SETUP);
  int rc = s->ops->mg_setup_jacobian(s->mg, s->uu, s->dt, NULL);
  PH_END(PFHIP_PROF_PH_PC_SETUP);
  CHK(rc);
  s->npe++;
  s->nnilpre = s->nni;
  s->precondcurrent = 1;
  return PFHIP_OK;
}

/* KinsolPC: zero-guess V-cycle; z = M^{-1} r (r, z may alias) */
static int psolve(pfhip_solver *s, const double *r, double *z)
{
  s->nps++;
  PH_BEGIN(PFHIP_PROF_PH_PC_SOLVE);
  int rc = PFHIP_OK;
  if (r == z) {
    rc = s->ops->vec_scale(s->pb, 1.0, r, s->vtemp2, NULL);
    if (!rc) rc = s->ops->mg_apply(s->mg, s->vtemp2, z, NULL);
  } else {
    rc = s->ops->mg_apply(s->mg, r, z, NULL);
  }
  PH_END(PFHIP_PROF_PH_PC_SOLVE);
  return rc;
}

/* z = M^{-1} (sc * r), r != z: the V-cycle applies the factor where it
 * reads r, as if r had been scaled first.  When z is the next J.v's
 * direction (the difference quotient, no overland), the V-cycle also
 * leaves that quotient's sums over uu and z on the device. */
static int psolve_scaled(pfhip_solver *s, const double *r, double sc, double *z)
{
  s->nps++;
  PH_BEGIN(PFHIP_PROF_PH_PC_SOLVE);
  int rc;
  s->dq_ready = NULL;
  if (s->ops->mg_apply_dq && s->ops->residual_jv_ready && !s->o.use_jacobian &&
      !s->pb->overland) {
    rc = s->ops->mg_apply_dq(s->mg, r, sc, z, s->uu, NULL);
    if (!rc) s->dq_ready = z;
  } else {
    rc = s->ops->mg_apply_scaled(s->mg, r, sc, z, NULL);
  }
  PH_END(PFHIP_PROF_PH_PC_SOLVE);
  return rc;
}

/* KINSpgmrAtimes: KINSolMatVec or difference quotient KS/kinspgmr.c:481-531 */
static int atimes_eval(pfhip_solver *s, const double *v, double *z)
{
  pfhip_problem *pb = s->pb;
  if (s->o.use_jacobian) {
    if (s->new_uu) {
      CHK(pfhip_richards_jacobian_eval(pb, s->uu, s->dt, 0, s->J, NULL));
      s->new_uu = 0;
    }
    return pfhip_matvec(pb, (const double *const *)s->J, 0, v, z, NULL);
  }
  if (!pb->overland && s->dq_ready == v && s->ops->residual_jv_ready) {
    /* the DQ sums came with v from the V-cycle */
    s->dq_ready = NULL;
    s->nfe++;
    return s->ops->residual_jv_ready(pb, s->uu, v, s->sqrt_relfunc, s->fval, z, s->cur_p_old,
                                     s->cur_s_old, s->dt, NULL);
  }
  if (!pb->overland && s->ops->residual_jv_dq) {
    /* the DQ sums, sigma and the quotient without leaving the device */
    s->nfe++;
    return s->ops->residual_jv_dq(pb, s->uu, v, s->sqrt_relfunc, s->fval, z, s->cur_p_old,
                                  s->cur_s_old, s->dt, NULL);
  }
  double sutsv, vtv, sq1norm;
  CHK(s->ops->vec_dq_sums(pb, s->uu, v, &sutsv, &vtv, &sq1norm, NULL));
  double sign = (sutsv >= 0.0) ? 1.0 : -1.0;
  double a = fabs(sutsv);
  double sigma = sign * s->sqrt_relfunc * ((a > sq1norm) ? a : sq1norm) / vtv;
  double sigma_inv = 1.0 / sigma;
  if (!pb->overland) {
    /* u + sigma v, F and the quotient in one residual launch */
    s->nfe++;
    return s->ops->residual_jv(pb, s->uu, v, sigma, s->fval, sigma_inv, z, s->cur_p_old,
                             s->cur_s_old, s->dt, NULL);
  }
  CHK(s->ops->vec_linear_sum(pb, 1.0, s->uu, sigma, v, s->vtemp1, NULL));
  CHK(func_eval(s, s->vtemp1, s->vtemp2));
  return s->ops->vec_linear_sum(pb, sigma_inv, s->vtemp2, -sigma_inv, s->fval, z, NULL);
}

static int atimes(pfhip_solver *s, const double *v, double *z)
{
  PH_BEGIN(PFHIP_PROF_PH_JV);
  int rc = atimes_eval(s, v, z);
  PH_END(PFHIP_PROF_PH_JV);
  return rc;
}

/* ModifiedGS (KS/iterativ.c:30-73), each step one fused axpy+dot pass */
static int modified_gs_eval(pfhip_solver *s, int k, double *new_vk_norm)
{
  pfhip_problem *pb = s->pb;
  double **v = s->V, **h = s->Hes;
  double vk2;
  CHK(s->ops->vec_dot(pb, v[k], v[k], &vk2, NULL));
  double vk_norm = sqrt(vk2);
  int km1 = k - 1;
  double d;
  CHK(pfhip_vec_axpy_dot(pb, 0.0, NULL, v[k], v[0], &d, NULL));
  h[0][km1] = d;
  for (int i = 1; i < k; i++) {
    CHK(pfhip_vec_axpy_dot(pb, h[i - 1][km1], v[i - 1], v[k], v[i], &d, NULL));
    h[i][km1] = d;
  }
  double n2;
  CHK(pfhip_vec_axpy_dot(pb, h[k - 1][km1], v[k - 1], v[k], v[k], &n2, NULL));
  *new_vk_norm = sqrt(n2);
  double temp = 1000.0 * vk_norm;
  if ((temp + (*new_vk_norm)) != temp) return PFHIP_OK;
  double new_norm_2 = 0.0;
  for (int i = 0; i < k; i++) {
    double np_;
    CHK(s->ops->vec_dot(pb, v[i], v[k], &np_, NULL));
    temp = 1000.0 * h[i][km1];
    if ((temp + np_) == temp) continue;
    h[i][km1] += np_;
    CHK(s->ops->vec_linear_sum(pb, 1.0, v[k], -np_, v[i], v[k], NULL));
    new_norm_2 += np_ * np_;
  }
  if (new_norm_2 != 0.0) {
    double t = (*new_vk_norm) * (*new_vk_norm) - new_norm_2;
    *new_vk_norm = (t > 0.0) ? sqrt(t) : 0.0;
  }
  return PFHIP_OK;
}

/* CGS2: two passes of block classical Gram-Schmidt against V[0..k-1]
 * (scaled: V[q] read as vsc[q] * V[q]) */
static int cgs2_eval(pfhip_solver *s, int k, double *new_vk_norm, int scaled)
{
  pfhip_problem *pb = s->pb;
  double *h1 = s->hbuf, *h2 = s->hbuf + (s->l_max + 2);
  int km1 = k - 1;
  double n2;
  if (scaled) {
    CHK(s->ops->cgs2_scaled(pb, (const double *const *)s->V, s->vsc, k, s->V[k], h1, h2, &n2,
                            NULL));
  } else if (s->ops->cgs2 && k <= PFHIP_SLOT_H2) {
    CHK(s->ops->cgs2(pb, (const double *const *)s->V, k, s->V[k], h1, h2, &n2, NULL));
  } else {
    CHK(s->ops->vec_multi_dot(pb, (const double *const *)s->V, k, s->V[k], h1, NULL));
    CHK(s->ops->vec_multi_axpy(pb, (const double *const *)s->V, k, h1, s->V[k], NULL, NULL));
    CHK(s->ops->vec_multi_dot(pb, (const double *const *)s->V, k, s->V[k], h2, NULL));
    CHK(s->ops->vec_multi_axpy(pb, (const double *const *)s->V, k, h2, s->V[k], &n2, NULL));
  }
  for (int i = 0; i < k; i++) s->Hes[i][km1] = h1[i] + h2[i];
  *new_vk_norm = sqrt(n2 > 0.0 ? n2 : 0.0);
  return PFHIP_OK;
}

static int modified_gs(pfhip_solver *s, int k, double *new_vk_norm)
{
  PH_BEGIN(PFHIP_PROF_PH_ORTHO);
  int rc = modified_gs_eval(s, k, new_vk_norm);
  PH_END(PFHIP_PROF_PH_ORTHO);
  return rc;
}

static int cgs2(pfhip_solver *s, int k, double *new_vk_norm, int scaled)
{
  PH_BEGIN(PFHIP_PROF_PH_ORTHO);
  int rc = cgs2_eval(s, k, new_vk_norm, scaled);
  PH_END(PFHIP_PROF_PH_ORTHO);
  return rc;
}

static int qrfact(int n, double **h, double *q)
{
  int nm1 = n - 1, code = 0;
  for (int k = 0; k < nm1; k++) {
    int i = 2 * k;
    double t1 = h[k][nm1], t2 = h[k + 1][nm1], c = q[i], sn = q[i + 1];
    h[k][nm1] = c * t1 - sn * t2;
    h[k + 1][nm1] = sn * t1 + c * t2;
  }
  double t1 = h[nm1][nm1], t2 = h[n][nm1], c, sn;
  if (t2 == 0.0) { c = 1.0; sn = 0.0; }
  else if (fabs(t2) >= fabs(t1)) {
    double t3 = t1 / t2;
    sn = -1.0 / sqrt(1.0 + t3 * t3);
    c = -sn * t3;
  } else {
    double t3 = t2 / t1;
    c = 1.0 / sqrt(1.0 + t3 * t3);
    sn = -c * t3;
  }
  q[2 * nm1] = c;
  q[2 * nm1 + 1] = sn;
  if ((h[nm1][nm1] = c * t1 - sn * t2) == 0.0) code = n;
  return code;
}

static int qrsol(int n, double **h, double *q, double *b)
{
  for (int k = 0; k < n; k++) {
    double c = q[2 * k], sn = q[2 * k + 1], t1 = b[k], t2 = b[k + 1];
    b[k] = c * t1 - sn * t2;
    b[k + 1] = sn * t1 + c * t2;
  }
  for (int k = n - 1; k >= 0; k--) {
    if (h[k][k] == 0.0) return k + 1;
    b[k] /= h[k][k];
    for (int i = 0; i < k; i++) b[i] -= b[k] * h[i][k];
  }
  return 0;
}

#define SPGMR_SUCCESS 0
#define SPGMR_RES_REDUCED 1
#define SPGMR_CONV_FAIL 2
#define SPGMR_QRFACT_FAIL 5

/* SpgmrSolve (KS/spgmr.c:168-450), right preconditioning, s1 = s2 = 1;
 * with krylov = 1 the flexible variant keeps Z[l] = M^{-1} V[l].  The
 * right-hand side is b = -fneg (the Newton system J p = -F): ||b|| is
 * ||fneg|| and V[0] = b / ||b|| is formed in one pass. */
/* fneg_norm: ||fneg|| when the caller already holds it (computed by the same
 * reduction), else < 0.
 * Flexible variant with a preconditioner: the basis is kept unnormalised,
 * V[k] = w_k with vsc[k] = 1/||w_k|| beside it (V[0] is fneg itself with
 * vsc[0] = -1/||fneg||), and every pass that reads V[k] (the V-cycle's
 * right-hand side, the Gram-Schmidt products) takes vsc[k] * V[k], the
 * value the scaling pass would have stored: the same bits without that
 * pass.  A restart first stores the scaled vectors. */
static int spgmr_solve_body(pfhip_solver *s, double *x, const double *fneg, double fneg_norm,
                            double delta, double *res_norm, int *nli, int *ret)
{
  pfhip_problem *pb = s->pb;
  double **V = s->V, **Hes = s->Hes, *vsc = s->vsc;
  int l_max = s->l_max, flex = (s->o.krylov == 1);
  int preOnRight = s->o.precond != 0;
  const int defer = flex && preOnRight && s->ops->cgs2_scaled && s->ops->mg_apply_scaled &&
                    l_max < PFHIP_SLOT_H2;
  double beta, rotation_product, r_norm, s_product, rho = 0;
  int krydim = 0, converged = 0;
  /* the flexible variant with the fused update leaves x unwritten until
   * its first update overwrites it; returns before that zero it */
  int x_unset = flex && s->ops->vec_combine_step;
#define X_ZERO_IF_UNSET() \
  do { if (x_unset) { CHK(s->ops->vec_const(pb, 0.0, x, NULL)); x_unset = 0; } } while (0)
  *nli = 0;
  if (fneg_norm >= 0.0) r_norm = fneg_norm;
  else CHK(l2norm(s, fneg, &r_norm));
  *res_norm = beta = r_norm;
  if (r_norm <= delta) { X_ZERO_IF_UNSET(); *ret = SPGMR_SUCCESS; return PFHIP_OK; }
  if (!flex) CHK(s->ops->vec_const(pb, 0.0, s->xcor, NULL));

  for (int ntries = 0; ntries <= s->max_restarts; ntries++) {
    for (int i = 0; i <= l_max; i++)
      for (int j = 0; j < l_max; j++) Hes[i][j] = 0.0;
    rotation_product = 1.0;
    for (int i = 0; i <= l_max; i++) vsc[i] = 1.0;
    if (ntries == 0 && defer) {
      V[0] = (double *)fneg;
      vsc[0] = -1.0 / r_norm;
    } else if (ntries == 0) {
      CHK(s->ops->vec_scale(pb, -1.0 / r_norm, fneg, V[0], NULL));
    } else {
      CHK(s->ops->vec_scale(pb, 1.0 / r_norm, V[0], V[0], NULL));
    }
    /* a restart cycle adds Z_c y_c to x: the first cycle's Arnoldi relation
     * no longer gives J x, so lsolve falls back to the difference quotient */
    if (ntries > 0) s->jp_ok = 0;
    int l;
    for (l = 0; l < l_max; l++) {
      (*nli)++;
      krydim = l + 1;
      const double *av = V[l];
      if (preOnRight) {
        double *z = flex ? s->Z[l] : s->vtemp;
        if (defer) CHK(psolve_scaled(s, V[l], vsc[l], z));
        else CHK(psolve(s, V[l], z));
        av = z;
      }
      CHK(atimes(s, av, V[l + 1]));
      if (flex) CHK(cgs2(s, l + 1, &Hes[l + 1][l], defer));
      else CHK(modified_gs(s, l + 1, &Hes[l + 1][l]));
      if (flex && ntries == 0)
        for (int i = 0; i <= l + 1; i++) s->Hraw[(size_t)i * l_max + l] = Hes[i][l];
      if (qrfact(krydim, Hes, s->givens) != 0) {
        X_ZERO_IF_UNSET();
        *ret = SPGMR_QRFACT_FAIL;
        return PFHIP_OK;
      }
      rotation_product *= s->givens[2 * l + 1];
      *res_norm = rho = fabs(rotation_product * r_norm);
      if (rho <= delta) { converged = 1; break; }
      if (defer) vsc[l + 1] = 1.0 / Hes[l + 1][l];
      else CHK(s->ops->vec_scale(pb, 1.0 / Hes[l + 1][l], V[l + 1], V[l + 1], NULL));
    }
    s->yg[0] = r_norm;
    for (int i = 1; i <= krydim; i++) s->yg[i] = 0.0;
    if (qrsol(krydim, Hes, s->givens, s->yg) != 0) {
      X_ZERO_IF_UNSET();
      *ret = SPGMR_QRFACT_FAIL;
      return PFHIP_OK;
    }
    if (flex) {
      /* x += Z y directly (flexible GMRES); without a preconditioner the
       * search directions are the basis vectors themselves (Z = V) */
      double **Zs = preOnRight ? s->Z : V;
      if (s->ops->vec_combine_step) {
        /* x = Z y in the first cycle (x not read), with the line search's
         * x.x and scaled step length from the same pass */
        CHK(s->ops->vec_combine_step(pb, (const double *const *)Zs, krydim, s->yg, x, s->uu,
                                     x_unset, &s->pp_dot, &s->pp_steplength, NULL));
        x_unset = 0;
        s->pp_stats_valid = 1;
      } else {
        CHK(s->ops->vec_combine(pb, (const double *const *)Zs, krydim, s->yg, x, NULL));
      }
      if (ntries == 0) {
        /* Arnoldi relation of this cycle: J Z = V Hbar, so with x = Z y
         * (x started at 0), J x = V g with g = Hbar y; V has orthonormal
         * columns and V[0] = -F / r_norm, hence ||J x|| = ||g|| and
         * F.(J x) = -r_norm g[0] */
        double g2 = 0.0, g0 = 0.0;
        for (int i = 0; i <= krydim; i++) {
          double gi = 0.0;
          for (int j = (i > 0 ? i - 1 : 0); j < krydim; j++)
            gi += s->Hraw[(size_t)i * l_max + j] * s->yg[j];
          if (i == 0) g0 = gi;
          g2 += gi * gi;
        }
        s->jp_ok = 1;
        s->jp_norm2 = g2;
        s->jp_fdot = -beta * g0;
      }
    } else {
      for (int k = 0; k < krydim; k++)
        CHK(s->ops->vec_linear_sum(pb, s->yg[k], V[k], 1.0, s->xcor, s->xcor, NULL));
    }
    if (converged) {
      if (!flex) {
        if (preOnRight) CHK(psolve(s, s->xcor, s->vtemp));
        else CHK(s->ops->vec_scale(pb, 1.0, s->xcor, s->vtemp, NULL));
        CHK(s->ops->vec_linear_sum(pb, 1.0, x, 1.0, s->vtemp, x, NULL));
      }
      *ret = SPGMR_SUCCESS;
      return PFHIP_OK;
    }
    if (ntries == s->max_restarts) break;
    s_product = 1.0;
    for (int i = krydim; i > 0; i--) {
      s->yg[i] = s_product * s->givens[2 * i - 2];
      s_product *= s->givens[2 * i - 1];
    }
    s->yg[0] = s_product;
    r_norm *= s_product;
    for (int i = 0; i <= krydim; i++) s->yg[i] *= r_norm;
    r_norm = fabs(r_norm);
    if (defer) {
      /* store the scaled basis (V[0] back in its own buffer) */
      for (int k = 0; k <= krydim; k++) {
        if (k == 0 && V[0] != s->V0_own) {
          CHK(s->ops->vec_scale(pb, vsc[0], V[0], s->V0_own, NULL));
          V[0] = s->V0_own;
        } else if (vsc[k] != 1.0) {
          CHK(s->ops->vec_scale(pb, vsc[k], V[k], V[k], NULL));
        }
        vsc[k] = 1.0;
      }
    }
    CHK(s->ops->vec_scale(pb, s->yg[0], V[0], V[0], NULL));
    for (int k = 1; k <= krydim; k++)
      CHK(s->ops->vec_linear_sum(pb, s->yg[k], V[k], 1.0, V[0], V[0], NULL));
  }
  if (rho < beta) {
    if (!flex) {
      if (preOnRight) CHK(psolve(s, s->xcor, s->vtemp));
      else CHK(s->ops->vec_scale(pb, 1.0, s->xcor, s->vtemp, NULL));
      CHK(s->ops->vec_linear_sum(pb, 1.0, x, 1.0, s->vtemp, x, NULL));
    }
    *ret = SPGMR_RES_REDUCED;
    return PFHIP_OK;
  }
  *ret = SPGMR_CONV_FAIL;
  return PFHIP_OK;
#undef X_ZERO_IF_UNSET
}

static int spgmr_solve(pfhip_solver *s, double *x, const double *fneg, double fneg_norm,
                       double delta, double *res_norm, int *nli, int *ret)
{
  int rc = spgmr_solve_body(s, x, fneg, fneg_norm, delta, res_norm, nli, ret);
  s->V[0] = s->V0_own;
  return rc;
}

/* KINSpgmrSolve KS/kinspgmr.c:344-420 for J xx = -fval; bb receives J xx;
 * *lret = KINSOL-level return */
static int lsolve(pfhip_solver *s, double *xx, double *bb, int *lret)
{
  pfhip_problem *pb = s->pb;
  int nli_inc, sret;
  double res_norm;
  s->pp_stats_valid = 0;
  s->dq_ready = NULL;
  /* the flexible variant's fused update writes xx without reading it */
  if (!(s->o.krylov == 1 && s->ops->vec_combine_step)) CHK(s->ops->vec_const(pb, 0.0, xx, NULL));
  s->new_uu = 1;
  s->jp_ok = 0;
  /* s->fnorm is l2norm(fval) of the current iterate: the line search (or
   * the initial check) computed it with the same reduction */
  CHK(spgmr_solve(s, xx, s->fval, s->fnorm, s->eps, &res_norm, &nli_inc, &sret));
  s->nli += nli_inc;
  if (sret != 0) s->ncfl++;
  if (s->o.krylov == 1 && s->jp_ok) {
    /* flexible variant: ||J p|| and F.(J p) from the Arnoldi relation of
     * the solve (no further residual evaluation); the reference evaluates
     * the difference quotient at p instead (KS/kinspgmr.c:400-412) */
    s->sJpnorm = sqrt(s->jp_norm2);
    s->sfdotJp = s->jp_fdot;
  } else {
    CHK(atimes(s, xx, bb));
    /* ||J p|| and F.(J p) from one pass: the (u.v, v.v) sums of the DQ
     * reduction are the same per-cell products in the same order as the
     * two separate dot products */
    double jp2, jp1;
    CHK(s->ops->vec_dq_sums(pb, s->fval, bb, &s->sfdotJp, &jp2, &jp1, NULL));
    s->sJpnorm = sqrt(jp2);
  }
  *lret = (sret == SPGMR_SUCCESS || sret == SPGMR_RES_REDUCED) ? 0 : 8;
  return PFHIP_OK;
}

/* KINLinSolDrv KS/kinsol.c:1424-1470 */
static int linsoldrv(pfhip_solver *s, double *bb, double *xx, int *ret)
{
  int setupNonNull = s->o.precond != 0;
  if (s->nni - s->nnilpre >= 1) s->pthrsh = 2.0;
  for (;;) {
    s->precondcurrent = 0;
    if ((s->pthrsh > 1.5) && setupNonNull) CHK(pc_setup(s));
    int lret;
    CHK(lsolve(s, xx, bb, &lret));
    if (lret != 1) { *ret = lret; return PFHIP_OK; }
    if (!setupNonNull || s->precondcurrent) { *ret = 8; return PFHIP_OK; }
    s->pthrsh = 2.0;
  }
}

static int eval_unew(pfhip_solver *s, double rl, double *fnormp, double *f1normp)
{
  s->stop_valid = 0;
  if (!s->pb->overland && s->ops->residual_step_stats) {
    /* unew = uu + rl pp formed inside the residual launch, which also
     * gathers KINStop's maxima; fval.fval from the same synchronisation */
    double ff;
    s->nfe++;
    PH_BEGIN(PFHIP_PROF_PH_NL_F_EVAL);
    int rc = s->ops->residual_step_stats(s->pb, s->uu, s->pp, rl, s->unew, s->fval,
                                         s->cur_p_old, s->cur_s_old, s->dt, &ff,
                                         &s->stop_fmax, &s->stop_steplength, NULL);
    PH_END(PFHIP_PROF_PH_NL_F_EVAL);
    CHK(rc);
    s->stop_valid = 1;
    *fnormp = sqrt(ff);
    *f1normp = 0.5 * (*fnormp) * (*fnormp);
    return PFHIP_OK;
  }
  if (!s->pb->overland) {
    /* unew = uu + rl pp formed inside the residual launch */
    s->nfe++;
    PH_BEGIN(PFHIP_PROF_PH_NL_F_EVAL);
    int rc = s->ops->residual_step(s->pb, s->uu, s->pp, rl, s->unew, s->fval, s->cur_p_old,
                                   s->cur_s_old, s->dt, NULL);
    PH_END(PFHIP_PROF_PH_NL_F_EVAL);
    CHK(rc);
  } else {
    CHK(s->ops->vec_linear_sum(s->pb, 1.0, s->uu, rl, s->pp, s->unew, NULL));
    CHK(func(s, s->unew, s->fval));
  }
  CHK(l2norm(s, s->fval, fnormp));
  *f1normp = 0.5 * (*fnormp) * (*fnormp);
  return PFHIP_OK;
}

/* KINLineSearch KS/kinsol.c:1228-1405 (no constraints) */
static int linesearch(pfhip_solver *s, double *fnormp, double *f1normp, int *maxStep,
                      int *gret)
{
  pfhip_problem *pb = s->pb;
  double pnorm, ratio = 1.0, alpha = 0.0001, beta = 0.9;
  double slpi, rlmin, rlength, rl, rlmax, rldiff, rltmp, rlprev = 0.0;
  double pt1trl, rllo, rlincr, alpha_cond, beta_cond = 0.0;
  *maxStep = 0;
  *gret = 0;
  int have_rlength = 0;
  if (s->pp_stats_valid) {
    /* gathered by the pass that formed pp, from the same uu */
    pnorm = sqrt(s->pp_dot);
    rlength = s->pp_steplength;
    have_rlength = 1;
  } else if (s->ops->vec_dot_scsteplength) {
    /* ||pp|| and the scaled step length from one pass; the step length is
     * taken again if pp is scaled down */
    double pp2;
    CHK(s->ops->vec_dot_scsteplength(pb, s->uu, s->pp, &pp2, &rlength, NULL));
    pnorm = sqrt(pp2);
    have_rlength = 1;
  } else {
    CHK(l2norm(s, s->pp, &pnorm));
  }
  if (pnorm > s->mxnewtstep) {
    ratio = s->mxnewtstep / pnorm;
    CHK(s->ops->vec_scale(pb, ratio, s->pp, s->pp, NULL));
    pnorm = s->mxnewtstep;
    have_rlength = 0;
  }
  s->stepl = pnorm;
  slpi = s->sfdotJp * ratio;
  if (!have_rlength) CHK(s->ops->vec_scsteplength(pb, s->uu, s->pp, &rlength, NULL));
  rlmin = s->scsteptol / rlength;
  rl = 1.0;
  int nfesav = s->nfe;
  for (;;) {
    CHK(eval_unew(s, rl, fnormp, f1normp));
    alpha_cond = s->f1norm + alpha * slpi * rl;
    if ((*f1normp) <= alpha_cond) break;
    if (rl < rlmin) {
      s->stop_valid = 0;
      CHK(s->ops->vec_scale(pb, 1.0, s->uu, s->unew, NULL));
      CHK(func(s, s->unew, s->fval));
      CHK(l2norm(s, s->fval, fnormp));
      *f1normp = 0.5 * (*fnormp) * (*fnormp);
      s->nbktrk = s->nfe - nfesav - 1;
      *gret = 1;
      return PFHIP_OK;
    }
    rltmp = -slpi / (2.0 * ((*f1normp) - s->f1norm - slpi));
    if (rltmp > 0.5 * rl) rltmp = 0.5 * rl;
    rlprev = rl;
    pt1trl = 0.1 * rl;
    rl = (pt1trl > rltmp) ? pt1trl : rltmp;
  }
  beta_cond = s->f1norm + beta * slpi * rl;
  if (*f1normp < beta_cond) {
    if (rl == 1.0 && pnorm < s->stepl) {
      rlmax = s->stepl / pnorm;
      do {
        rlprev = rl;
        rl = (2.0 * rl < rlmax) ? 2.0 * rl : rlmax;
        CHK(eval_unew(s, rl, fnormp, f1normp));
        alpha_cond = s->f1norm + alpha * slpi * rl;
        beta_cond = s->f1norm + beta * slpi * rl;
      } while ((*f1normp) <= alpha_cond && (*f1normp) < beta_cond && (rl < rlmax));
    }
    if (rl < 1.0 || ((rl > 1.0) && (*f1normp > alpha_cond))) {
      rllo = (rl < rlprev) ? rl : rlprev;
      rldiff = fabs(rlprev - rl);
      do {
        rlincr = 0.5 * rldiff;
        rl = rllo + rlincr;
        CHK(eval_unew(s, rl, fnormp, f1normp));
        alpha_cond = s->f1norm + alpha * slpi * rl;
        beta_cond = s->f1norm + beta * slpi * rl;
        if ((*f1normp) > alpha_cond) rldiff = rlincr;
        else if (*f1normp < beta_cond) {
          rllo = rl;
          rldiff = rldiff - rlincr;
        }
      } while (*f1normp > alpha_cond || ((*f1normp < beta_cond) && (rldiff > rlmin)));
      if ((*f1normp) < beta_cond) {
        CHK(eval_unew(s, rllo, fnormp, f1normp));
        s->nbcf++;
      }
    }
  }
  s->nbktrk = s->nfe - nfesav - 1;
  s->sfdotJp = s->sfdotJp * rl * ratio;
  s->sJpnorm = s->sJpnorm * rl * ratio;
  if (rl * pnorm > 0.99 * s->mxnewtstep) *maxStep = 1;
  return PFHIP_OK;
}

/* KINInexactNewton KS/kinsol.c:1152-1220 */
static int inexact_newton(pfhip_solver *s, double *fnormp, double *f1normp, int *maxStep)
{
  double pnorm, ratio = 1.0;
  *maxStep = 0;
  if (s->pp_stats_valid) pnorm = sqrt(s->pp_dot);
  else CHK(l2norm(s, s->pp, &pnorm));
  if (pnorm > s->mxnewtstep) {
    ratio = s->mxnewtstep / pnorm;
    CHK(s->ops->vec_scale(s->pb, ratio, s->pp, s->pp, NULL));
    pnorm = s->mxnewtstep;
  }
  s->stepl = pnorm;
  s->sfdotJp *= ratio;
  s->sJpnorm *= ratio;
  CHK(eval_unew(s, 1.0, fnormp, f1normp));
  if (pnorm > 0.99 * s->mxnewtstep) *maxStep = 1;
  return PFHIP_OK;
}

/* KINForcingTerm KS/kinsol.c:1054-1090; etaflag 1 = Walker1, 2 = Walker2 */
static void forcing_term(pfhip_solver *s, double fnormp)
{
  double eta_max = 0.9, eta_min = 0.0001, eta_safe = 0.5;
  if (s->etaflag == 1) {
    double lm = sqrt(s->fnorm * s->fnorm + 2.0 * s->sfdotJp + s->sJpnorm * s->sJpnorm);
    eta_safe = (s->eta <= 0.0) ? 0.0 : pow(s->eta, s->ealpha);
    s->eta = fabs(fnormp - lm) / s->fnorm;
  }
  if (s->etaflag == 2) {
    eta_safe = s->egamma * ((s->eta <= 0.0) ? 0.0 : pow(s->eta, s->ealpha));
    double r = fnormp / s->fnorm;
    s->eta = s->egamma * ((r <= 0.0) ? 0.0 : pow(r, s->ealpha));
  }
  if (eta_safe < 0.1) eta_safe = 0.0;
  if (s->eta < eta_safe) s->eta = eta_safe;
  if (s->eta < eta_min) s->eta = eta_min;
  if (s->eta > eta_max) s->eta = eta_max;
}

/* KINStop KS/kinsol.c:1507-1580 */
static int kin_stop(pfhip_solver *s, int maxStepTaken, int gret, int *ret)
{
  int setupNonNull = s->o.precond != 0;
  *ret = 0;
  if (gret == 1) {
    if (setupNonNull && !s->precondcurrent) { s->pthrsh = 2.0; return PFHIP_OK; }
    *ret = 3;
    return PFHIP_OK;
  }
  double fmax, rlength;
  if (s->stop_valid) {
    /* gathered by the residual launch of the accepted point */
    fmax = s->stop_fmax;
    rlength = s->stop_steplength;
  } else {
    /* max|fval| and the scaled step length from one pass */
    CHK(s->ops->kin_stop_norms(s->pb, s->fval, s->unew, s->uu, &fmax, &rlength, NULL));
  }
  if (fmax <= s->fnormtol) { *ret = 1; return PFHIP_OK; }
  if (rlength <= s->scsteptol) {
    if (!s->precondcurrent) { s->pthrsh = 2.0; return PFHIP_OK; }
    *ret = 2;
    return PFHIP_OK;
  }
  if (s->nni >= s->mxiter) { *ret = 4; return PFHIP_OK; }
  if (maxStepTaken) s->ncscmx++;
  else s->ncscmx = 0;
  if (s->ncscmx == 5) { *ret = 5; return PFHIP_OK; }
  s->pthrsh = rlength;
  return PFHIP_OK;
}

static double unit_roundoff(void)
{
  double u = 1.0;
  volatile double one_plus_u = 1.0 + u;
  while (one_plus_u != 1.0) { u /= 2.0; one_plus_u = 1.0 + u; }
  return u * 2.0;
}

static int kinsol_solve(pfhip_solver *s, double *p, const double *p_old,
                        const double *s_old, double dt, double time, pfhip_kin_stats *st);

int pfhip_kinsol_nonlin_solver(pfhip_solver *s, double *p, const double *p_old,
                               const double *s_old, double dt, double time,
                               pfhip_kin_stats *st)
{
  PH_BEGIN(PFHIP_PROF_PH_KINSOL);
  int rc = kinsol_solve(s, p, p_old, s_old, dt, time, st);
  PH_END(PFHIP_PROF_PH_KINSOL);
  return rc;
}

static int kinsol_solve(pfhip_solver *s, double *p, const double *p_old,
                        const double *s_old, double dt, double time, pfhip_kin_stats *st)
{
  pfhip_problem *pb = s->pb;
  const pfhip_solver_opts *o = &s->o;
  CHK(s->ops->bc_pressure(pb, time, NULL));
  s->uu = p;
  s->unew = s->unew_own;
  s->cur_p_old = p_old;
  s->cur_s_old = s_old;
  s->dt = dt;
  s->nni = s->nnilpre = s->nfe = s->nbcf = s->nbktrk = s->ncscmx = 0;
  s->nli = s->nps = s->npe = s->ncfl = 0;

  /* KINSolInit */
  s->uround = unit_roundoff();
  s->fnormtol = o->residual_tol > 0.0 ? o->residual_tol : pow(s->uround, 1.0 / 3.0);
  s->scsteptol = o->step_tol > 0.0 ? o->step_tol : pow(s->uround, 2.0 / 3.0);
  s->mxiter = o->max_iter > 0 ? o->max_iter : 200;
  s->sqrt_relfunc = (o->deriv_eps > 0.0) ? sqrt(o->deriv_eps) : sqrt(s->uround);
  /* eta setup, KS/kinsol.c:526-575 */
  if (o->eta_choice == 0) {
    s->etaflag = 0;
    s->callForcingTerm = 0;
    s->eta = o->eta_value;
  } else if (o->eta_choice == 1) {
    s->etaflag = 1;
    s->callForcingTerm = 1;
    s->ealpha = (1.0 + sqrt(5.0)) / 2.0;
    s->eta = 0.5;
  } else {
    s->etaflag = 2;
    s->callForcingTerm = 1;
    s->ealpha = o->eta_alpha;
    s->egamma = o->eta_gamma;
    if (s->ealpha <= 1.0 || s->ealpha > 2.0) s->ealpha = 2.0;
    if (s->egamma <= 0.0 || s->egamma > 1.0) s->egamma = 0.9;
    s->eta = 0.5;
  }
  if (s->eta <= 0.0 || s->eta > 1.0) s->eta = 0.1;

  int ret = 0;
  /* KINInitialStop */
  CHK(s->ops->vec_scale(pb, 1.0, p, s->unew, NULL));
  CHK(func(s, p, s->fval));
  /* mxnewtstep from ||u|| (KS/kinsol.c:518-522), then max |fval| and
   * ||fval|| (KS/kinsol.c:640-644); the three reductions are order-fixed,
   * so one pass gives the values of three */
  double fmax, un, ff;
  if (s->ops->kinsol_init_norms) {
    CHK(s->ops->kinsol_init_norms(pb, p, s->fval, &fmax, &ff, &un, NULL));
    un = sqrt(un);
  } else {
    CHK(l2norm(s, p, &un));
    CHK(s->ops->vec_max_norm(pb, s->fval, &fmax, NULL));
    ff = -1.0;
  }
  s->mxnewtstep = 1000.0 * un;
  if (s->mxnewtstep < 1.0) s->mxnewtstep = 1.0;
  if (fmax <= 0.01 * s->fnormtol) {
    ret = 2;
  } else {
    if (ff >= 0.0) s->fnorm = sqrt(ff);
    else CHK(l2norm(s, s->fval, &s->fnorm));
    s->f1norm = 0.5 * s->fnorm * s->fnorm;
    s->pthrsh = 2.0;
    double epsmin = 0.01 * s->fnormtol;
    for (;;) {
      s->nni++;
      s->stop_valid = 0;
      s->eps = (s->eta + s->uround) * s->fnorm;
      if (s->eps < epsmin) s->eps = epsmin;
      CHK(linsoldrv(s, s->unew, s->pp, &ret));
      if (ret != 0) break;
      double fnormp = 0, f1normp = 0;
      int maxStep = 0, gret = 0;
      if (o->globalization == 0) CHK(inexact_newton(s, &fnormp, &f1normp, &maxStep));
      else CHK(linesearch(s, &fnormp, &f1normp, &maxStep, &gret));
      if (s->nbcf > 10) { ret = 7; break; }
      if (s->callForcingTerm) forcing_term(s, fnormp);
      s->fnorm = fnormp;
      CHK(kin_stop(s, maxStep, gret, &ret));
      /* uu = unew by exchanging the two buffers instead of copying */
      { double *t = s->uu; s->uu = s->unew; s->unew = t; }
      s->f1norm = f1normp;
      if (ret != 0) break;
    }
  }
  if (s->uu != p) {
    /* the accepted iterate is in the solver's own buffer: hand it back in p */
    CHK(s->ops->vec_scale(pb, 1.0, s->uu, p, NULL));
    s->uu = p;
  }
  s->unew = s->unew_own;
  CHK(s->ops->synchronize());
  if (st) {
    st->ret = ret;
    st->nni = s->nni; st->nli = s->nli; st->nfe = s->nfe; st->npe = s->npe;
    st->nps = s->nps; st->nbcf = s->nbcf; st->nbktrk = s->nbktrk;
    st->ncfl = s->ncfl; st->fnorm = s->fnorm;
  }
  if (ret == 1 || ret == 2) return PFHIP_OK;
  pfhip_set_error("KINSOL returned %d", ret);
  return PFHIP_ERR_SOLVER;
}

int pfhip_advance_richards(pfhip_solver *s, double *p, double *sat, double t0,
                           double dt, int nsteps, int max_conv_failures, int *newton,
                           int *gmres, int *fevals, double *dt_used, int *steps_done)
{
  pfhip_problem *pb = s->pb;
  double t = t0;
  long bytes = s->ng * (long)sizeof(double);
  if (steps_done) *steps_done = 0;
  CHK(s->ops->saturation(pb, p, sat, PFHIP_CALCFCN, NULL));
  for (int n = 0; n < nsteps; n++) {
    int converged = 1, failures = 0, tn = 0, tl = 0, tf = 0;
    double dtc = dt;
    do {
      if (converged) {
        CHK(s->ops->vec_scale(pb, 1.0, sat, s->s_old, NULL));
        CHK(s->ops->vec_scale(pb, 1.0, p, s->p_old, NULL));
      } else {
        t = t - dtc;
        dtc = 0.5 * dtc;
        CHK(s->ops->vec_scale(pb, 1.0, s->s_old, sat, NULL));
        CHK(s->ops->vec_scale(pb, 1.0, s->p_old, p, NULL));
      }
      t += dtc;
      pfhip_kin_stats st;
      int rc = pfhip_kinsol_nonlin_solver(s, p, s->p_old, s->s_old, dtc, t, &st);
      if (rc != PFHIP_OK && rc != PFHIP_ERR_SOLVER) return rc;
      tn += st.nni; tl += st.nli; tf += st.nfe;
      if (rc == PFHIP_ERR_SOLVER) { converged = 0; failures++; }
      else converged = 1;
    } while (!converged && failures < max_conv_failures);
    if (!converged) {
      pfhip_set_error("time step %d failed after %d convergence failures", n, failures);
      return PFHIP_ERR_SOLVER;
    }
    CHK(s->ops->saturation(pb, p, sat, PFHIP_CALCFCN, NULL));
    if (newton) newton[n] = tn;
    if (gmres) gmres[n] = tl;
    if (fevals) fevals[n] = tf;
    if (dt_used) dt_used[n] = dtc;
    if (steps_done) *steps_done = n + 1;
  }
  (void)bytes;
  return s->ops->synchronize();
}

/* ------------------------------------------------------------------------
 * Time-step control.  TIME_EPSILON (PL/general.h:257) and the 1000
 * subdivisions of a base unit (PL/time_cycle_data.c:150).
 * ---------------------------------------------------------------------- */
#include <float.h>
#define TS_TIME_EPSILON (FLT_EPSILON * 10)
#define TS_SUBDIVISIONS 1000

/* pfround, PL/general.h:176 */
static int ts_round(double x) { return (x < 0.0) ? (int)(x - 0.5) : (int)(x + 0.5); }

double pfhip_next_transition(const pfhip_bc_spec *bc, int nbc, double base_unit,
                             double start_time, double stop_time, double time)
{
  double deltat = -1.0;
  int assigned = 0;
  for (int c = 0; c < nbc; c++) {
    const pfhip_bc_spec *b = &bc[c];
    int division = b->n_intervals;
    if (division < 1 || division > PFHIP_MAX_INTERVALS) continue;
    int cycle_length = 0;
    for (int i = 0; i < division; i++) cycle_length += b->lengths[i];
    if (cycle_length <= 0) continue;
    double time_defined = (b->repeat > 0)
                              ? ((double)(b->repeat * cycle_length * base_unit))
                              : stop_time;
    if (!(time < time_defined && division > 1)) continue;
    /* time in whole base units, rounded at 1/1000 of a unit first */
    int dtime = ts_round(time / (base_unit / TS_SUBDIVISIONS)) / TS_SUBDIVISIONS;
    int dstart = ts_round(start_time / (base_unit / TS_SUBDIVISIONS)) / TS_SUBDIVISIONS;
    int n = (dtime - dstart) / cycle_length;
    int completed = (dtime - dstart - 1) % cycle_length;
    int in = 0, total = b->lengths[0];
    while (completed >= total && in + 1 < division) {
      in++;
      total += b->lengths[in];
    }
    int next = (in + 1) % division;
    double transition;
    if (next == 0) {
      n++;
      transition = start_time + ((double)(n * cycle_length)) * base_unit;
    } else {
      transition = start_time + ((double)(n * cycle_length + total)) * base_unit;
    }
    if (transition < time_defined) {
      if (assigned) {
        double d = transition - time;
        deltat = (deltat < d) ? deltat : d;
      } else {
        deltat = transition - time;
        assigned = 1;
      }
    }
  }
  /* a transition closer than TIME_EPSILON does not count */
  double test = time + deltat;
  double diff = test - time;
  if (diff <= TS_TIME_EPSILON) deltat = -1.0;
  return deltat;
}

int pfhip_select_time_step(const pfhip_time_control *tc, const pfhip_bc_spec *bc, int nbc,
                           double time, double *dt, char *dt_info)
{
  const pfhip_time_step *ts = &tc->ts;
  if (ts->type == PFHIP_TS_CONSTANT) {
    *dt = ts->value;
  } else if (ts->type == PFHIP_TS_GROWTH) {
    if (*dt == 0.0) {
      *dt = ts->initial_step;
    } else {
      *dt = *dt * ts->factor;
      if (*dt < ts->min_step) *dt = ts->min_step;
      if (*dt > ts->max_step) *dt = ts->max_step;
    }
  } else {
    pfhip_set_error("select_time_step: unknown TimeStep type %d", ts->type);
    return PFHIP_ERR_ARG;
  }
  double bc_dt = pfhip_next_transition(bc, nbc, tc->base_unit, tc->start_time, tc->stop_time,
                                       time);
  if (bc_dt <= 0.0) bc_dt = *dt;
  if (*dt > bc_dt) {
    *dt = bc_dt;
    *dt_info = 'b';
  } else {
    *dt_info = 'p';
  }
  return PFHIP_OK;
}

/* The surface predictor on the step's initial guess.  The reference uses
 * the face velocities and surface fluxes its last residual evaluation left:
 * none (zeros) before the run's first solve, else those of the previous
 * step's accepted pressure, which are recomputed here from p with the BC
 * values of that solve.  (After a failed attempt the reference's come from
 * the failed solve's last iterate; here they are those of the step's
 * initial pressure.) */
static int surface_predict(pfhip_solver *s, double *p, const double *sat, int have_vel,
                           double dt_vel, double dt, const pfhip_time_control *tc)
{
  pfhip_problem *pb = s->pb;
  const int nx = pb->d.nx, ny = pb->d.ny, nz = pb->d.nz;
  const long n[3] = { (long)(nx + 1) * ny * nz, (long)nx * (ny + 1) * nz,
                      (long)nx * ny * (nz + 1) };
  for (int q = 0; q < 3; q++)
    if (!s->vel[q]) CHK(pfhip_dev_alloc(&s->vel[q], n[q]));
  const double *qx = NULL, *qy = NULL;
  if (!have_vel) {
    for (int q = 0; q < 3; q++) CHK(pfhip_zero(s->vel[q], n[q], NULL));
  } else {
    CHK(pfhip_face_velocity(pb, p, p, dt_vel, s->vel[0], s->vel[1], s->vel[2], NULL));
    if (pb->overland) { qx = pb->q2[0]; qy = pb->q2[1]; }
  }
  return pfhip_surface_predictor(pb, p, sat, s->vel[0], s->vel[1], s->vel[2], qx, qy, dt,
                                 tc->surface_predictor_pressure, tc->surface_lateral_flows,
                                 NULL);
}

int pfhip_advance_richards_ts(pfhip_solver *s, double *p, double *sat,
                              const pfhip_time_control *tc, pfhip_time_state *state,
                              int max_steps, int max_conv_failures, pfhip_step_log *log,
                              int *steps_done)
{
  pfhip_problem *pb = s->pb;
  const pfhip_bc_spec *bc = pb->d.bc;
  if (steps_done) *steps_done = 0;
  if (tc->dump_interval < 0.0 && (int)(-tc->dump_interval) < 1) {
    pfhip_set_error("advance_richards_ts: DumpInterval %g is not a whole step count",
                    tc->dump_interval);
    return PFHIP_ERR_ARG;
  }
  if (state->dump_index < 1) state->dump_index = 1;
  CHK(s->ops->saturation(pb, p, sat, PFHIP_CALCFCN, NULL));
  int done = 0;
  while (done < max_steps && state->iteration_number < tc->max_iterations &&
         state->t < tc->stop_time) {
    double t = state->t, dt = state->dt;
    char info = 'p';
    int converged = 1, failures = 0, dump = 0, tn = 0, tl = 0, tf = 0;
    do {
      if (converged) {
        CHK(pfhip_select_time_step(tc, bc, 6, t, &dt, &info));
        CHK(s->ops->vec_scale(pb, 1.0, sat, s->s_old, NULL));
        CHK(s->ops->vec_scale(pb, 1.0, p, s->p_old, NULL));
      } else {
        t = t - dt;
        double new_dt = 0.5 * dt;
        double test_time = t + new_dt;
        if (test_time - t > TS_TIME_EPSILON) {
          dt = new_dt;
        } else {
          pfhip_set_error("Time increment is too small; solver has failed (t = %g)", t);
          return PFHIP_ERR_SOLVER;
        }
        CHK(s->ops->vec_scale(pb, 1.0, s->s_old, sat, NULL));
        CHK(s->ops->vec_scale(pb, 1.0, s->p_old, p, NULL));
      }
      /* dump times (PL/solver_richards.c:3141-3176) */
      dump = 0;
      if (tc->dump_interval > 0.0) {
        double print_dt = tc->start_time + state->dump_index * tc->dump_interval - t;
        if ((dt + TS_TIME_EPSILON) > print_dt) {
          if (fabs(dt - print_dt) > TS_TIME_EPSILON) dt = print_dt;
          info = 'p';
          dump = 1;
        }
      } else if (tc->dump_interval < 0.0) {
        if ((state->iteration_number % (-(int)tc->dump_interval)) == 0) dump = 1;
      }
      /* last step (:3210-3232) */
      if ((t + dt) >= tc->stop_time) {
        dt = tc->stop_time - t;
        info = 'f';
      }
      t += dt;
      if (tc->surface_predictor) {
        int have_vel = state->iteration_number > 0 || failures > 0;
        CHK(surface_predict(s, p, sat, have_vel, state->dt > 0.0 ? state->dt : dt, dt, tc));
      }
      pfhip_kin_stats st;
      int rc = pfhip_kinsol_nonlin_solver(s, p, s->p_old, s->s_old, dt, t, &st);
      if (rc != PFHIP_OK && rc != PFHIP_ERR_SOLVER) return rc;
      tn += st.nni; tl += st.nli; tf += st.nfe;
      if (rc == PFHIP_ERR_SOLVER) { converged = 0; failures++; }
      else converged = 1;
    } while (!converged && failures < max_conv_failures);
    if (!converged) {
      pfhip_set_error("Time step failed for time %12.4e after %d convergence failures", t,
                      failures);
      return PFHIP_ERR_SOLVER;
    }
    CHK(s->ops->saturation(pb, p, sat, PFHIP_CALCFCN, NULL));
    state->t = t;
    state->dt = dt;
    state->iteration_number++;
    if (dump) state->dump_index++;
    if (log) {
      pfhip_step_log *r = &log[done];
      r->time = t; r->dt = dt; r->dt_info = info; r->dumped = dump;
      r->newton = tn; r->gmres = tl; r->fevals = tf; r->conv_failures = failures;
    }
    done++;
    if (steps_done) *steps_done = done;
  }
  return s->ops->synchronize();
}
