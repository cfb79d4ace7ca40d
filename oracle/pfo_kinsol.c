/*
 * pfo_kinsol.c -- oracle (test infrastructure only) restatement of the
 * embedded KINSOL inexact-Newton solver with SPGMR that ParFlow builds when
 * SUNDIALS is absent, and of the AdvanceRichards per-step contract.
 *
 *   KINSolInit           KS/kinsol.c:363-650
 *   KINSol main loop     KS/kinsol.c:690-860
 *   KINForcingTerm       KS/kinsol.c:1054-1090
 *   KINInitialStop       KS/kinsol.c:1128-1137
 *   KINInexactNewton     KS/kinsol.c:1152-1220
 *   KINLineSearch        KS/kinsol.c:1228-1405
 *   KINLinSolDrv         KS/kinsol.c:1424-1470
 *   KINStop              KS/kinsol.c:1507-1580
 *   KINSpgmrSolve/Atimes KS/kinspgmr.c:344-531
 *   SpgmrSolve           KS/spgmr.c:168-450
 *   ModifiedGS/QR        KS/iterativ.c:30-286
 *   KinsolPC / callbacks PL/kinsol_pc.c:58-178, PL/kinsol_nonlin_solver.c:144-236
 *   AdvanceRichards      PL/solver_richards.c:2970-3400, 3661-3669
 * uscale = fscale = 1 (PL/kinsol_nonlin_solver.c:700-712).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include "pfo.h"

#define IDX(i, j, k) ((i) + 1 + ((j) + 1) * sy + ((k) + 1) * sz)

typedef struct {
  pfo_problem *pb;
  const pfo_solver_opts *o;
  const double *p_old, *s_old;
  double dt;
  long ng;
  /* KINSOL memory */
  double *uu, *unew, *fval, *pp, *vtemp1, *vtemp2;
  double fnorm, f1norm, eta, ealpha, egamma, sJpnorm, sfdotJp, stepl;
  double mxnewtstep, sqrt_relfunc, uround, fnormtol, scsteptol, eps;
  double pthrsh;
  int etaflag, callForcingTerm, nni, nnilpre, nfe, nbcf, nbktrk, ncscmx;
  int precondcurrent, mxiter;
  /* SPGMR */
  int l_max, max_restarts;
  double **V, **Hes, *givens, *xcor, *yg, *vtemp;
  int nli, nps, npe, ncfl;
  /* Jacobian / PC */
  double *J[7], *P[7];
  int new_uu;
  pfo_mg *mg;
  double *pcbuf;
} kin_t;

/* --- interior-only N_Vector ops (PL/vector_utilities.c) --------------- */
#define FORALL(pb) \
  for (int k = 0; k < (pb)->nz; k++) \
    for (int j = 0; j < (pb)->ny; j++) \
      for (int i = 0; i < (pb)->nx; i++)

static void v_linsum(const pfo_problem *pb, double a, const double *x,
                     double b, const double *y, double *z)
{
  const int sy = pb->nx + 2, sz = (pb->nx + 2) * (pb->ny + 2);
  FORALL(pb) { int ip = IDX(i, j, k); z[ip] = a * x[ip] + b * y[ip]; }
}
static void v_scale(const pfo_problem *pb, double c, const double *x, double *z)
{
  const int sy = pb->nx + 2, sz = (pb->nx + 2) * (pb->ny + 2);
  FORALL(pb) { int ip = IDX(i, j, k); z[ip] = c * x[ip]; }
}
static void v_const(const pfo_problem *pb, double c, double *z)
{
  const int sy = pb->nx + 2, sz = (pb->nx + 2) * (pb->ny + 2);
  FORALL(pb) { int ip = IDX(i, j, k); z[ip] = c; }
}
/* KINScSteplength with uscale = 1: max |s|/(1+|u|) */
static double v_scsteplength(const pfo_problem *pb, const double *u,
                             const double *s)
{
  const int sy = pb->nx + 2, sz = (pb->nx + 2) * (pb->ny + 2);
  double m = 0.0;
  FORALL(pb)
  {
    int ip = IDX(i, j, k);
    double v = fabs(s[ip] / (1.0 + fabs(u[ip])));
    if (v > m) m = v;
  }
  return m;
}

static void func(kin_t *K, const double *u, double *f)
{
  pfo_nl_function_eval(K->pb, u, K->p_old, K->s_old, K->dt, f);
  K->nfe++;
}

/* KINSolInitPC: PC matrix at uu + preconditioner setup */
static void pc_setup(kin_t *K)
{
  pfo_jacobian_eval(K->pb, K->uu, K->dt, !K->o->pc_matrix, K->P);
  pfo_mg_setup(K->mg, K->P);
  K->npe++;
  K->nnilpre = K->nni;
  K->precondcurrent = 1;
}

/* KINSpgmrPSolve -> KinsolPC: zero-guess V-cycle, in place */
static void psolve(kin_t *K, double *z)
{
  memcpy(K->pcbuf, z, sizeof(double) * K->ng);
  pfo_mg_apply(K->mg, K->pcbuf, z);
  K->nps++;
}

/* KINSpgmrAtimes: analytic (KINSolMatVec) or difference quotient */
static void atimes(kin_t *K, const double *v, double *z)
{
  pfo_problem *pb = K->pb;
  if (K->o->use_jacobian)
  {
    if (K->new_uu)
    {
      pfo_jacobian_eval(pb, K->uu, K->dt, 0, K->J);
      K->new_uu = 0;
    }
    pfo_matvec(pb, K->J, 0, v, z);
    return;
  }
  /* KS/kinspgmr.c:481-531 */
  double sutsv = pfo_dot(pb, K->uu, v);
  double vtv = pfo_dot(pb, v, v);
  double sq1norm = pfo_l1norm(pb, v);
  double sign = (sutsv >= 0.0) ? 1.0 : -1.0;
  double sutsv_abs = fabs(sutsv);
  double sigma = sign * K->sqrt_relfunc
                 * ((sutsv_abs > sq1norm) ? sutsv_abs : sq1norm) / vtv;
  double sigma_inv = 1.0 / sigma;
  v_linsum(pb, 1.0, K->uu, sigma, v, K->vtemp1);
  func(K, K->vtemp1, K->vtemp2);
  v_linsum(pb, sigma_inv, K->vtemp2, -sigma_inv, K->fval, z);
}

/* ModifiedGS KS/iterativ.c:30-73 (p = l_max, i0 = 0) */
static void modified_gs(kin_t *K, int kk, double *new_vk_norm)
{
  pfo_problem *pb = K->pb;
  double **v = K->V, **h = K->Hes;
  double vk_norm = sqrt(pfo_dot(pb, v[kk], v[kk]));
  int km1 = kk - 1;
  for (int i = 0; i < kk; i++)
  {
    h[i][km1] = pfo_dot(pb, v[i], v[kk]);
    v_linsum(pb, 1.0, v[kk], -h[i][km1], v[i], v[kk]);
  }
  *new_vk_norm = sqrt(pfo_dot(pb, v[kk], v[kk]));
  double temp = 1000.0 * vk_norm;
  if ((temp + (*new_vk_norm)) != temp) return;
  double new_norm_2 = 0.0;
  for (int i = 0; i < kk; i++)
  {
    double new_product = pfo_dot(pb, v[i], v[kk]);
    temp = 1000.0 * h[i][km1];
    if ((temp + new_product) == temp) continue;
    h[i][km1] += new_product;
    v_linsum(pb, 1.0, v[kk], -new_product, v[i], v[kk]);
    new_norm_2 += new_product * new_product;
  }
  if (new_norm_2 != 0.0)
  {
    double np2 = (*new_vk_norm) * (*new_vk_norm) - new_norm_2;
    *new_vk_norm = (np2 > 0.0) ? sqrt(np2) : 0.0;
  }
}

/* QRfact job != 0 branch, KS/iterativ.c:190-232 */
static int qrfact(int n, double **h, double *q)
{
  int nm1 = n - 1, code = 0;
  for (int k = 0; k < nm1; k++)
  {
    int i = 2 * k;
    double t1 = h[k][nm1], t2 = h[k + 1][nm1];
    double c = q[i], s = q[i + 1];
    h[k][nm1] = c * t1 - s * t2;
    h[k + 1][nm1] = s * t1 + c * t2;
  }
  double t1 = h[nm1][nm1], t2 = h[n][nm1], c, s;
  if (t2 == 0.0) { c = 1.0; s = 0.0; }
  else if (fabs(t2) >= fabs(t1))
  {
    double t3 = t1 / t2;
    s = -1.0 / sqrt(1.0 + t3 * t3);
    c = -s * t3;
  }
  else
  {
    double t3 = t2 / t1;
    c = 1.0 / sqrt(1.0 + t3 * t3);
    s = -c * t3;
  }
  q[2 * nm1] = c;
  q[2 * nm1 + 1] = s;
  if ((h[nm1][nm1] = c * t1 - s * t2) == 0.0) code = n;
  return code;
}

/* QRsol KS/iterativ.c:253-286 */
static int qrsol(int n, double **h, double *q, double *b)
{
  int code = 0;
  for (int k = 0; k < n; k++)
  {
    double c = q[2 * k], s = q[2 * k + 1];
    double t1 = b[k], t2 = b[k + 1];
    b[k] = c * t1 - s * t2;
    b[k + 1] = s * t1 + c * t2;
  }
  for (int k = n - 1; k >= 0; k--)
  {
    if (h[k][k] == 0.0) { code = k + 1; break; }
    b[k] /= h[k][k];
    for (int i = 0; i < k; i++) b[i] -= b[k] * h[i][k];
  }
  return code;
}

#define SPGMR_SUCCESS 0
#define SPGMR_RES_REDUCED 1
#define SPGMR_CONV_FAIL 2
#define SPGMR_QRFACT_FAIL 5

/* SpgmrSolve with right preconditioning, s1 = s2 = 1 (KS/spgmr.c:168-450) */
static int spgmr_solve(kin_t *K, double *x, const double *b, double delta,
                       double *res_norm, int *nli, int *nps)
{
  pfo_problem *pb = K->pb;
  double **V = K->V, **Hes = K->Hes;
  int l_max = K->l_max;
  int preOnRight = K->o->precond != 0;
  double beta, rotation_product, r_norm, s_product, rho = 0;
  int krydim = 0, converged = 0;
  int nps0 = K->nps;
  *nli = 0;

  /* x is zero on entry (KINSpgmrSolve) */
  v_scale(pb, 1.0, b, K->vtemp);
  v_scale(pb, 1.0, K->vtemp, V[0]);
  v_scale(pb, 1.0, V[0], K->vtemp);
  v_scale(pb, 1.0, K->vtemp, V[0]);
  *res_norm = r_norm = beta = sqrt(pfo_dot(pb, V[0], V[0]));
  if (r_norm <= delta) { *nps = 0; return SPGMR_SUCCESS; }
  v_const(pb, 0.0, K->xcor);

  for (int ntries = 0; ntries <= K->max_restarts; ntries++)
  {
    for (int i = 0; i <= l_max; i++)
      for (int j = 0; j < l_max; j++) Hes[i][j] = 0.0;
    rotation_product = 1.0;
    v_scale(pb, 1.0 / r_norm, V[0], V[0]);
    int l;
    for (l = 0; l < l_max; l++)
    {
      (*nli)++;
      krydim = l + 1;
      v_scale(pb, 1.0, V[l], K->vtemp);
      if (preOnRight)
      {
        v_scale(pb, 1.0, K->vtemp, V[l + 1]);
        v_scale(pb, 1.0, V[l + 1], K->vtemp);
        psolve(K, K->vtemp);
      }
      atimes(K, K->vtemp, V[l + 1]);
      v_scale(pb, 1.0, V[l + 1], K->vtemp);
      v_scale(pb, 1.0, K->vtemp, V[l + 1]);
      modified_gs(K, l + 1, &(Hes[l + 1][l]));
      if (qrfact(krydim, Hes, K->givens) != 0) return SPGMR_QRFACT_FAIL;
      rotation_product *= K->givens[2 * l + 1];
      *res_norm = rho = fabs(rotation_product * r_norm);
      if (rho <= delta) { converged = 1; break; }
      v_scale(pb, 1.0 / Hes[l + 1][l], V[l + 1], V[l + 1]);
    }
    K->yg[0] = r_norm;
    for (int i = 1; i <= krydim; i++) K->yg[i] = 0.0;
    if (qrsol(krydim, Hes, K->givens, K->yg) != 0) return SPGMR_QRFACT_FAIL;
    for (int kx = 0; kx < krydim; kx++)
      v_linsum(pb, K->yg[kx], V[kx], 1.0, K->xcor, K->xcor);
    if (converged)
    {
      if (preOnRight) { v_scale(pb, 1.0, K->xcor, K->vtemp); psolve(K, K->vtemp); }
      else v_scale(pb, 1.0, K->xcor, K->vtemp);
      v_linsum(pb, 1.0, x, 1.0, K->vtemp, x);
      *nps = K->nps - nps0;
      return SPGMR_SUCCESS;
    }
    if (ntries == K->max_restarts) break;
    s_product = 1.0;
    for (int i = krydim; i > 0; i--)
    {
      K->yg[i] = s_product * K->givens[2 * i - 2];
      s_product *= K->givens[2 * i - 1];
    }
    K->yg[0] = s_product;
    r_norm *= s_product;
    for (int i = 0; i <= krydim; i++) K->yg[i] *= r_norm;
    r_norm = fabs(r_norm);
    v_scale(pb, K->yg[0], V[0], V[0]);
    for (int kx = 1; kx <= krydim; kx++)
      v_linsum(pb, K->yg[kx], V[kx], 1.0, V[0], V[0]);
  }
  if (rho < beta)
  {
    if (preOnRight) { v_scale(pb, 1.0, K->xcor, K->vtemp); psolve(K, K->vtemp); }
    else v_scale(pb, 1.0, K->xcor, K->vtemp);
    v_linsum(pb, 1.0, x, 1.0, K->vtemp, x);
    *nps = K->nps - nps0;
    return SPGMR_RES_REDUCED;
  }
  *nps = K->nps - nps0;
  return SPGMR_CONV_FAIL;
}

/* KINSpgmrSolve KS/kinspgmr.c:344-420 */
static int lsolve(kin_t *K, double *xx, double *bb, double *res_norm)
{
  pfo_problem *pb = K->pb;
  int nli_inc, nps_inc;
  v_const(pb, 0.0, xx);
  K->new_uu = 1;
  int ret = spgmr_solve(K, xx, bb, K->eps, res_norm, &nli_inc, &nps_inc);
  K->nli += nli_inc;
  if (ret != 0) K->ncfl++;
  atimes(K, xx, bb);
  K->sJpnorm = pfo_l2norm(pb, bb);
  K->sfdotJp = pfo_dot(pb, K->fval, bb);
  if (ret == SPGMR_SUCCESS || ret == SPGMR_RES_REDUCED) return 0;
  return 8; /* KINSOL_KRYLOV_FAILURE */
}

/* KINLinSolDrv KS/kinsol.c:1424-1470 */
static int linsoldrv(kin_t *K, double *bb, double *xx)
{
  pfo_problem *pb = K->pb;
  int setupNonNull = K->o->precond != 0;
  if (K->nni - K->nnilpre >= 1) K->pthrsh = 2.0;
  for (;;)
  {
    K->precondcurrent = 0;
    if ((K->pthrsh > 1.5) && setupNonNull)
      pc_setup(K);
    v_scale(pb, -1.0, K->fval, bb);
    double res_norm;
    int ret = lsolve(K, xx, bb, &res_norm);
    if (ret != 1) return ret;
    if (!setupNonNull) return 8;
    if (K->precondcurrent) return 8;
    K->pthrsh = 2.0;
  }
}

/* KINLineSearch KS/kinsol.c:1228-1405 (no constraints) */
static int linesearch(kin_t *K, double *fnormp, double *f1normp, int *maxStep)
{
  pfo_problem *pb = K->pb;
  double pnorm, ratio = 1.0, alpha = 0.0001, beta = 0.9;
  double slpi, rlmin, rlength, rl, rlmax, rldiff, rltmp, rlprev = 0.0;
  double pt1trl, rllo, rlincr, alpha_cond, beta_cond;
  int ivio = 0;
  *maxStep = 0;
  pnorm = pfo_l2norm(pb, K->pp);
  if (pnorm > K->mxnewtstep)
  {
    ratio = K->mxnewtstep / pnorm;
    v_scale(pb, ratio, K->pp, K->pp);
    pnorm = K->mxnewtstep;
  }
  K->stepl = pnorm;
  slpi = K->sfdotJp * ratio;
  rlength = v_scsteplength(pb, K->uu, K->pp);
  rlmin = K->scsteptol / rlength;
  rl = 1.0;
  int nfesav = K->nfe;
  for (;;)
  {
    v_linsum(pb, 1.0, K->uu, rl, K->pp, K->unew);
    func(K, K->unew, K->fval);
    *fnormp = pfo_l2norm(pb, K->fval);
    *f1normp = 0.5 * (*fnormp) * (*fnormp);
    alpha_cond = K->f1norm + alpha * slpi * rl;
    if ((*f1normp) <= alpha_cond) break;
    if (rl < rlmin)
    {
      v_scale(pb, 1.0, K->uu, K->unew);
      func(K, K->unew, K->fval);
      *fnormp = pfo_l2norm(pb, K->fval);
      *f1normp = 0.5 * (*fnormp) * (*fnormp);
      K->nbktrk = K->nfe - nfesav - 1;
      return 1;
    }
    rltmp = -slpi / (2.0 * ((*f1normp) - K->f1norm - slpi));
    if (rltmp > 0.5 * rl) rltmp = 0.5 * rl;
    rlprev = rl;
    pt1trl = 0.1 * rl;
    rl = (pt1trl > rltmp) ? pt1trl : rltmp;
  }
  beta_cond = K->f1norm + beta * slpi * rl;
  if (*f1normp < beta_cond)
  {
    if (rl == 1.0 && pnorm < K->stepl)
    {
      rlmax = K->stepl / pnorm;
      if (ivio == 1) rlmax = 1.0;
      do
      {
        rlprev = rl;
        rl = (2.0 * rl < rlmax) ? 2.0 * rl : rlmax;
        v_linsum(pb, 1.0, K->uu, rl, K->pp, K->unew);
        func(K, K->unew, K->fval);
        *fnormp = pfo_l2norm(pb, K->fval);
        *f1normp = 0.5 * (*fnormp) * (*fnormp);
        alpha_cond = K->f1norm + alpha * slpi * rl;
        beta_cond = K->f1norm + beta * slpi * rl;
      } while ((*f1normp) <= alpha_cond && (*f1normp) < beta_cond
               && (rl < rlmax));
    }
    if (rl < 1.0 || ((rl > 1.0) && (*f1normp > alpha_cond)))
    {
      rllo = (rl < rlprev) ? rl : rlprev;
      rldiff = fabs(rlprev - rl);
      do
      {
        rlincr = 0.5 * rldiff;
        rl = rllo + rlincr;
        v_linsum(pb, 1.0, K->uu, rl, K->pp, K->unew);
        func(K, K->unew, K->fval);
        *fnormp = pfo_l2norm(pb, K->fval);
        *f1normp = 0.5 * (*fnormp) * (*fnormp);
        alpha_cond = K->f1norm + alpha * slpi * rl;
        beta_cond = K->f1norm + beta * slpi * rl;
        if ((*f1normp) > alpha_cond) rldiff = rlincr;
        else if (*f1normp < beta_cond)
        {
          rllo = rl;
          rldiff = rldiff - rlincr;
        }
      } while (*f1normp > alpha_cond ||
               ((*f1normp < beta_cond) && (rldiff > rlmin)));
      if ((*f1normp) < beta_cond)
      {
        v_linsum(pb, 1.0, K->uu, rllo, K->pp, K->unew);
        func(K, K->unew, K->fval);
        *fnormp = pfo_l2norm(pb, K->fval);
        *f1normp = 0.5 * (*fnormp) * (*fnormp);
        K->nbcf++;
      }
    }
  }
  K->nbktrk = K->nfe - nfesav - 1;
  K->sfdotJp = K->sfdotJp * rl * ratio;
  K->sJpnorm = K->sJpnorm * rl * ratio;
  if (rl * pnorm > 0.99 * K->mxnewtstep) *maxStep = 1;
  return 0;
}

/* KINInexactNewton KS/kinsol.c:1152-1220 (no constraints) */
static int inexact_newton(kin_t *K, double *fnormp, double *f1normp,
                          int *maxStep)
{
  pfo_problem *pb = K->pb;
  double pnorm, ratio = 1.0;
  *maxStep = 0;
  pnorm = pfo_l2norm(pb, K->pp);
  if (pnorm > K->mxnewtstep)
  {
    ratio = K->mxnewtstep / pnorm;
    v_scale(pb, ratio, K->pp, K->pp);
    pnorm = K->mxnewtstep;
  }
  K->stepl = pnorm;
  K->sfdotJp = K->sfdotJp * ratio;
  K->sJpnorm = K->sJpnorm * ratio;
  v_linsum(pb, 1.0, K->uu, 1.0, K->pp, K->unew);
  func(K, K->unew, K->fval);
  *fnormp = pfo_l2norm(pb, K->fval);
  *f1normp = 0.5 * (*fnormp) * (*fnormp);
  if (pnorm > 0.99 * K->mxnewtstep) *maxStep = 1;
  return 0;
}

/* KINForcingTerm KS/kinsol.c:1054-1090 */
static void forcing_term(kin_t *K, double fnormp)
{
  double eta_max = 0.9, eta_min = 0.0001, eta_safe = 0.5, linmodel_norm;
  if (K->etaflag == 1)
  {
    linmodel_norm = sqrt(K->fnorm * K->fnorm + 2.0 * K->sfdotJp
                         + K->sJpnorm * K->sJpnorm);
    eta_safe = (K->eta <= 0.0) ? 0.0 : pow(K->eta, K->ealpha);
    K->eta = fabs(fnormp - linmodel_norm) / K->fnorm;
  }
  if (K->etaflag == 2)
  {
    eta_safe = K->egamma * ((K->eta <= 0.0) ? 0.0 : pow(K->eta, K->ealpha));
    double r = fnormp / K->fnorm;
    K->eta = K->egamma * ((r <= 0.0) ? 0.0 : pow(r, K->ealpha));
  }
  if (eta_safe < 0.1) eta_safe = 0.0;
  if (K->eta < eta_safe) K->eta = eta_safe;
  if (K->eta < eta_min) K->eta = eta_min;
  if (K->eta > eta_max) K->eta = eta_max;
}

/* KINStop KS/kinsol.c:1507-1580 */
static int kin_stop(kin_t *K, int maxStepTaken, int globalstratret)
{
  pfo_problem *pb = K->pb;
  int setupNonNull = K->o->precond != 0;
  if (globalstratret == 1)
  {
    if (setupNonNull && !K->precondcurrent)
    {
      K->pthrsh = 2.0;
      return 0;
    }
    return 3;
  }
  double fmax = pfo_maxnorm(pb, K->fval);
  if (fmax <= K->fnormtol) return 1;
  v_linsum(pb, 1.0, K->unew, -1.0, K->uu, K->vtemp1);
  double rlength = v_scsteplength(pb, K->unew, K->vtemp1);
  if (rlength <= K->scsteptol)
  {
    if (!K->precondcurrent)
    {
      K->pthrsh = 2.0;
      return 0;
    }
    return 2;
  }
  if (K->nni >= K->mxiter) return 4;
  if (maxStepTaken) K->ncscmx++;
  else K->ncscmx = 0;
  if (K->ncscmx == 5) return 5;
  K->pthrsh = rlength;
  return 0;
}

static double unit_roundoff(void)
{
  double u = 1.0;
  volatile double one_plus_u = 1.0 + u;
  while (one_plus_u != 1.0)
  {
    u /= 2.0;
    one_plus_u = 1.0 + u;
  }
  return u * 2.0;
}

/* Returns 0 when ParFlow would treat the solve as converged
 * (KINSOL_SUCCESS or KINSOL_INITIAL_GUESS_OK, PL/kinsol_nonlin_solver.c
 * :442-445), nonzero otherwise; st->ret is the raw KINSOL code. */
int pfo_kinsol(pfo_problem *pb, const pfo_solver_opts *opts, double *p,
               const double *p_old, const double *s_old, double dt,
               double time, pfo_kin_stats *st)
{
  kin_t K;
  memset(&K, 0, sizeof(K));
  K.pb = pb;
  K.o = opts;
  K.p_old = p_old;
  K.s_old = s_old;
  K.dt = dt;
  const long ng = (long)(pb->nx + 2) * (pb->ny + 2) * (pb->nz + 2);
  K.ng = ng;
  pfo_bc_pressure(pb, time);

  double *blk = (double*)calloc(ng * 8, sizeof(double));
  K.uu = p;
  K.unew = blk;
  K.fval = blk + ng;
  K.pp = blk + 2 * ng;
  K.vtemp1 = blk + 3 * ng;
  K.vtemp2 = blk + 4 * ng;
  K.xcor = blk + 5 * ng;
  K.vtemp = blk + 6 * ng;
  K.pcbuf = blk + 7 * ng;
  memcpy(K.unew, p, sizeof(double) * ng);

  K.l_max = opts->krylov_dim;
  K.max_restarts = (opts->max_restarts <= 0) ? 0 : opts->max_restarts;
  K.V = (double**)malloc(sizeof(double*) * (K.l_max + 1));
  for (int i = 0; i <= K.l_max; i++) K.V[i] = (double*)calloc(ng, sizeof(double));
  K.Hes = (double**)malloc(sizeof(double*) * (K.l_max + 1));
  for (int i = 0; i <= K.l_max; i++) K.Hes[i] = (double*)calloc(K.l_max, sizeof(double));
  K.givens = (double*)calloc(2 * K.l_max, sizeof(double));
  K.yg = (double*)calloc(K.l_max + 1, sizeof(double));
  for (int s = 0; s < 7; s++)
  {
    K.J[s] = opts->use_jacobian ? (double*)calloc(ng, sizeof(double)) : NULL;
    K.P[s] = (double*)calloc(ng, sizeof(double));
  }
  if (opts->precond)
    K.mg = pfo_mg_create_full(pb->nx, pb->ny, pb->nz, pb->dx, pb->dy, pb->dz,
                              opts->mg_max_levels, opts->mg_pre, opts->mg_post,
                              opts->mg_symmetric, opts->pc_matrix);

  /* KINSolInit */
  K.uround = unit_roundoff();
  K.fnormtol = opts->residual_tol;
  K.scsteptol = opts->step_tol;
  if (K.scsteptol <= 0.0) K.scsteptol = pow(K.uround, 2.0 / 3.0);
  if (K.fnormtol <= 0.0) K.fnormtol = pow(K.uround, 1.0 / 3.0);
  K.mxiter = (opts->max_iter > 0) ? opts->max_iter : 200;
  K.sqrt_relfunc = (opts->deriv_eps > 0.0) ? sqrt(opts->deriv_eps)
                                           : sqrt(K.uround);
  K.mxnewtstep = 1000.0 * pfo_l2norm(pb, p);
  if (K.mxnewtstep < 1.0) K.mxnewtstep = 1.0;
  K.etaflag = opts->eta_choice == 0 ? 2 /* ETACONSTANT */ :
              (opts->eta_choice == 1 ? 0 : 1);
  /* map to KINSOL's enum: ETACHOICE1 = 0, ETACHOICE2 = 1, ETACONSTANT = 2 */
  {
    int e = K.etaflag;
    K.callForcingTerm = (e != 2);
    if (e == 0) K.ealpha = (1.0 + sqrt(5.0)) / 2.0;
    if (e == 1) { K.ealpha = 2.0; K.egamma = 0.9; }
    K.eta = (e == 2) ? 0.1 : 0.5;
    if (e == 1)
    {
      K.ealpha = opts->eta_alpha;
      K.egamma = opts->eta_gamma;
      if (K.ealpha <= 1.0 || K.ealpha > 2.0) K.ealpha = 2.0;
      if (K.egamma <= 0.0 || K.egamma > 1.0) K.egamma = 0.9;
    }
    else if (e == 2)
      K.eta = opts->eta_value;
    if (K.eta <= 0.0 || K.eta > 1.0) K.eta = 0.1;
    /* forcing_term() uses 1 = choice 1 and 2 = choice 2 */
    K.etaflag = (e == 0) ? 1 : (e == 1) ? 2 : 0;
  }
  K.nfe = K.nni = K.nbcf = K.nbktrk = 0;

  int ret;
  /* KINInitialStop */
  func(&K, K.uu, K.fval);
  if (pfo_maxnorm(pb, K.fval) <= 0.01 * K.fnormtol)
  {
    ret = 2; /* KINSOL_INITIAL_GUESS_OK */
    goto done;
  }
  K.fnorm = pfo_l2norm(pb, K.fval);
  K.f1norm = 0.5 * K.fnorm * K.fnorm;

  K.ncscmx = 0;
  K.pthrsh = 2.0;
  double epsmin = 0.01 * K.fnormtol;
  for (;;)
  {
    K.nni++;
    K.eps = (K.eta + K.uround) * K.fnorm;
    if (K.eps < epsmin) K.eps = epsmin;
    ret = linsoldrv(&K, K.unew, K.pp);
    if (ret != 0) break;
    double fnormp = 0, f1normp = 0;
    int maxStep = 0, gret;
    if (opts->globalization == 0)
      gret = inexact_newton(&K, &fnormp, &f1normp, &maxStep);
    else
      gret = linesearch(&K, &fnormp, &f1normp, &maxStep);
    if (K.nbcf > 10) { ret = 7; break; }
    if (K.callForcingTerm) forcing_term(&K, fnormp);
    K.fnorm = fnormp;
    ret = kin_stop(&K, maxStep, gret);
    v_scale(pb, 1.0, K.unew, K.uu);
    K.f1norm = f1normp;
    if (ret != 0) break;
  }

done:
  if (st)
  {
    st->ret = ret;
    st->nni = K.nni; st->nli = K.nli; st->nfe = K.nfe; st->npe = K.npe;
    st->nps = K.nps; st->nbcf = K.nbcf; st->nbktrk = K.nbktrk;
    st->ncfl = K.ncfl; st->fnorm = K.fnorm;
  }
  for (int i = 0; i <= K.l_max; i++) { free(K.V[i]); free(K.Hes[i]); }
  free(K.V); free(K.Hes); free(K.givens); free(K.yg);
  for (int s = 0; s < 7; s++) { free(K.J[s]); free(K.P[s]); }
  if (K.mg) pfo_mg_destroy(K.mg);
  free(blk);
  return (ret == 1 || ret == 2) ? 0 : 1;
}

/* AdvanceRichards per-step contract (PL/solver_richards.c:2970-3400,
 * 3661-3669): copy old state, t += dt, Newton solve, on failure halve dt
 * and restore (up to MaxConvergenceFailures), then recompute S.
 * p_hist/s_hist (optional) receive nsteps+1 ghosted snapshots. */
int pfo_advance(pfo_problem *pb, const pfo_solver_opts *opts, double *p,
                double *s, double t0, double dt, int nsteps,
                int max_conv_failures, double *p_hist, double *s_hist,
                pfo_run_log *log)
{
  const long ng = (long)(pb->nx + 2) * (pb->ny + 2) * (pb->nz + 2);
  double *p_old = (double*)malloc(sizeof(double) * ng);
  double *s_old = (double*)malloc(sizeof(double) * ng);
  double t = t0;
  pfo_saturation(pb, p, s, PFO_CALCFCN);
  if (p_hist) memcpy(p_hist, p, sizeof(double) * ng);
  if (s_hist) memcpy(s_hist, s, sizeof(double) * ng);
  int status = 0;
  for (int n = 0; n < nsteps; n++)
  {
    int converged = 1, conv_failures = 0;
    double dtc = dt;
    int tot_nni = 0, tot_nli = 0, tot_nfe = 0;
    do
    {
      if (converged)
      {
        memcpy(s_old, s, sizeof(double) * ng);
        memcpy(p_old, p, sizeof(double) * ng);
      }
      else
      {
        t = t - dtc;
        dtc = 0.5 * dtc;
        memcpy(s, s_old, sizeof(double) * ng);
        memcpy(p, p_old, sizeof(double) * ng);
      }
      t += dtc;
      pfo_kin_stats st;
      int r = pfo_kinsol(pb, opts, p, p_old, s_old, dtc, t, &st);
      tot_nni += st.nni; tot_nli += st.nli; tot_nfe += st.nfe;
      if (r != 0) { converged = 0; conv_failures++; }
      else converged = 1;
    } while (!converged && conv_failures < max_conv_failures);
    if (!converged) { status = 1; break; }
    pfo_saturation(pb, p, s, PFO_CALCFCN);
    if (log && n < 4096)
    {
      log->newton[n] = tot_nni;
      log->gmres[n] = tot_nli;
      log->fevals[n] = tot_nfe;
      log->dt_used[n] = dtc;
      log->nsteps_done = n + 1;
    }
    if (p_hist) memcpy(p_hist + (n + 1) * ng, p, sizeof(double) * ng);
    if (s_hist) memcpy(s_hist + (n + 1) * ng, s, sizeof(double) * ng);
  }
  free(p_old);
  free(s_old);
  return status;
}
