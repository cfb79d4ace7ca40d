/*
 * ensemble.c -- many independent Richards problems advanced together.
 *
 * Every member runs the unchanged nonlinear driver of kinsol_host.c
 * (pfhip_advance_richards: AdvanceRichards PL/solver_richards.c:2970-3400
 * around KINSOL) in a coroutine of its own.  The driver reaches the device
 * only through its operations table; the ensemble's table records each
 * call as a request and switches back to the scheduler.  Once every live
 * member has a request pending, the scheduler serves them by kind with one
 * batched launch each (all vector operations and reductions together, all
 * residual forms, all PC setups, all V-cycles), copies the reduction
 * results back with one transfer and one synchronisation, and resumes the
 * members it served.  PC setups and V-cycles are held back while lighter
 * requests are pending, which keeps the members in step.  A member whose steps are done leaves the rotation: the set of
 * live members is the convergence mask of every launch.
 *
 * Members diverge freely (different Newton and GMRES counts, line-search
 * backtracks, dt halving); a round simply holds requests of several kinds.
 * One thread runs the scheduler and all coroutines (each on a stack of
 * its own, switched by pfhip_ens_switch), so the driver's scalar logic,
 * error buffer and call order are exactly those of a single-problem solve.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfhip_internal.h"

#define ENS_STACK (1L << 20)            /* coroutine stack (lazily touched) */
#define ENS_STAGE (64L << 20)           /* staging ring for member tables   */
#define ENS_QMAX 16                     /* requests a member queues a round */

/* Coroutine switch (System V x86-64): push the callee-saved registers,
 * store the stack pointer through the first argument, load the second and
 * pop the other side's registers.  No signal-mask system call per switch
 * (ucontext's swapcontext makes one), which matters at ~10^3 switches per
 * scheduler round. */
#if !defined(__x86_64__)
#error "ensemble.c: the coroutine switch is written for x86-64 hosts"
#endif
__asm__(".text\n"
        ".globl pfhip_ens_switch\n"
        ".hidden pfhip_ens_switch\n"
        ".type pfhip_ens_switch,@function\n"
        "pfhip_ens_switch:\n"
        "  pushq %rbp\n"
        "  pushq %rbx\n"
        "  pushq %r12\n"
        "  pushq %r13\n"
        "  pushq %r14\n"
        "  pushq %r15\n"
        "  movq %rsp, (%rdi)\n"
        "  movq %rsi, %rsp\n"
        "  popq %r15\n"
        "  popq %r14\n"
        "  popq %r13\n"
        "  popq %r12\n"
        "  popq %rbx\n"
        "  popq %rbp\n"
        "  ret\n"
        ".size pfhip_ens_switch, .-pfhip_ens_switch\n");
void pfhip_ens_switch(void **save_sp, void *load_sp);

typedef struct ens_member {
  pfhip_solver *s;
  pfhip_problem *pb;
  pfhip_mg *mg;
  double *scratch[4];
  void *sp;                             /* saved stack pointer             */
  void *stack;
  int live, rc;
  char err[256];                        /* message of the member's failure */
  int waiting;                          /* queue awaits the scheduler      */
  int qlen, nexec, rc_round;            /* queued / run this round / error */
  pfhip_ens_req q[ENS_QMAX];
  /* advance arguments and outputs */
  double *p, *sat, t0, dt;
  int nsteps, max_fail, steps_done;
  int *newton, *gmres, *fevals;
  double *dt_used;
} ens_member;

struct pfhip_ensemble {
  int n;
  ens_member *m;
  void *sched_sp;
  int cur;
  double *d_res, *h_res;
  double *d_part;                       /* block partials of vector ops    */
  pfhip_ens_stage st;
  pfhip_ens_req **lists[PFHIP_EOP_COUNT];
  int nlist[PFHIP_EOP_COUNT];
  pfhip_ens_req **vec, **outs;
  pfhip_ens_mgplan *mgplan;
  long rounds, launches;
};

/* the ensemble being advanced on this thread */
static __thread pfhip_ensemble *g_ens;

#define CHK(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

/* ---- the operations table of a member -------------------------------- */
static int is_heavy(int op) { return op == PFHIP_EOP_MG_SETUP || op == PFHIP_EOP_MG_APPLY; }

static pfhip_ens_req *new_req(int op)
{
  ens_member *m = &g_ens->m[g_ens->cur];
  pfhip_ens_req *r = &m->q[m->qlen];
  /* the fields every kind reads; term and result arrays are filled up to
   * nq / nout by the caller */
  r->mode = r->base = r->norm = r->nq = r->nout = r->rc = 0;
  r->cst = r->sig = r->fsc = r->dt = 0.0;
  r->x = r->y = r->u = NULL;
  r->z = r->z2 = NULL;
  r->xq[0] = r->xq[1] = NULL;
  r->op = op;
  r->member = g_ens->cur;
  r->pb = m->pb;
  r->mg = m->mg;
  for (int q = 0; q < 4; q++) r->scratch[q] = m->scratch[q];
  return r;
}

/* A request that returns nothing to the host and is light (a linear
 * combination without its norm, a residual form) only joins the member's
 * queue, and the member runs on; a request with results, a heavy one or a
 * full queue hands the queue to the scheduler.  The return code is the
 * first error of the queue's execution. */
static int submit(void)
{
  ens_member *m = &g_ens->m[g_ens->cur];
  pfhip_ens_req *r = &m->q[m->qlen++];
  if (r->nout == 0 && !is_heavy(r->op) && m->qlen < ENS_QMAX) return PFHIP_OK;
  pfhip_ens_switch(&m->sp, g_ens->sched_sp);
  return m->rc_round;
}

/* hands a non-empty queue to the scheduler (at a synchronisation point
 * and when the member ends) */
static int flush(void)
{
  ens_member *m = &g_ens->m[g_ens->cur];
  if (m->qlen == 0) return PFHIP_OK;
  pfhip_ens_switch(&m->sp, g_ens->sched_sp);
  return m->rc_round;
}

static int e_vec_scale(pfhip_problem *pb, double c, const double *x, double *z, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_LIN);
  r->nq = 1; r->c[0] = c; r->xq[0] = x; r->z = z;
  return submit();
}

static int e_vec_linear_sum(pfhip_problem *pb, double a, const double *x, double b,
                            const double *y, double *z, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_LIN);
  r->nq = 2; r->c[0] = a; r->xq[0] = x; r->c[1] = b; r->xq[1] = y; r->z = z;
  return submit();
}

static int e_vec_const(pfhip_problem *pb, double c, double *z, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_LIN);
  r->nq = 0; r->cst = c; r->z = z;
  return submit();
}

static int e_vec_dot(pfhip_problem *pb, const double *x, const double *y, double *out, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_DOT);
  r->x = x; r->y = y; r->out[0] = out; r->nout = 1;
  return submit();
}

static int e_vec_max_norm(pfhip_problem *pb, const double *x, double *out, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_MAXABS);
  r->x = x; r->out[0] = out; r->nout = 1;
  return submit();
}

static int e_vec_scsteplength(pfhip_problem *pb, const double *u, const double *s,
                              double *out, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_SCSTEP);
  r->x = u; r->y = s; r->out[0] = out; r->nout = 1;
  return submit();
}

static int e_vec_dq_sums(pfhip_problem *pb, const double *u, const double *v, double *a,
                         double *b, double *c, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_DQ);
  r->x = u; r->y = v; r->out[0] = a; r->out[1] = b; r->out[2] = c; r->nout = 3;
  return submit();
}

static int e_vec_multi_dot(pfhip_problem *pb, const double *const *V, int m, const double *w,
                           double *h, void *st)
{
  if (m < 0 || m > PFHIP_ENS_NRES) {
    pfhip_set_error("ensemble multi_dot over %d vectors (at most %d)", m, PFHIP_ENS_NRES);
    return PFHIP_ERR_ARG;
  }
  pfhip_ens_req *r = new_req(PFHIP_EOP_MDOT);
  r->nq = m; r->y = w;
  for (int q = 0; q < m; q++) { r->xq[q] = V[q]; r->out[q] = &h[q]; }
  r->nout = m;
  return submit();
}

static int lin_multi(const double *const *V, int m, const double *coef, double sign,
                     double *w, double *norm2)
{
  if (m < 0 || m > PFHIP_ENS_MAXQ) {
    pfhip_set_error("ensemble combination of %d vectors (at most %d)", m, PFHIP_ENS_MAXQ);
    return PFHIP_ERR_ARG;
  }
  pfhip_ens_req *r = new_req(PFHIP_EOP_LIN);
  r->base = 1; r->nq = m; r->z = w;
  for (int q = 0; q < m; q++) { r->c[q] = sign * coef[q]; r->xq[q] = V[q]; }
  if (norm2) { r->norm = 1; r->out[0] = norm2; r->nout = 1; }
  return submit();
}

static int e_vec_multi_axpy(pfhip_problem *pb, const double *const *V, int m, const double *h,
                            double *w, double *norm2, void *st)
{
  return lin_multi(V, m, h, -1.0, w, norm2);
}

static int e_vec_combine(pfhip_problem *pb, const double *const *V, int m, const double *y,
                         double *x, void *st)
{
  return lin_multi(V, m, y, 1.0, x, NULL);
}

static int e_kin_stop_norms(pfhip_problem *pb, const double *fval, const double *unew,
                            const double *uu, double *fmax, double *steplen, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_STOP);
  r->x = fval; r->y = unew; r->u = uu;
  r->out[0] = fmax; r->out[1] = steplen; r->nout = 2;
  return submit();
}

static int e_nl_function_eval(pfhip_problem *pb, const double *p, double *F,
                              const double *p_old, const double *s_old, double dt, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_RESIDUAL);
  r->mode = 0; r->x = p; r->z = F; r->xq[0] = p_old; r->xq[1] = s_old; r->dt = dt;
  return submit();
}

static int e_residual_jv(pfhip_problem *pb, const double *u, const double *v, double sigma,
                         const double *fval, double sinv, double *z, const double *p_old,
                         const double *s_old, double dt, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_RESIDUAL);
  r->mode = 1; r->x = u; r->y = v; r->u = fval; r->sig = sigma; r->fsc = sinv; r->z = z;
  r->xq[0] = p_old; r->xq[1] = s_old; r->dt = dt;
  return submit();
}

static int e_residual_step(pfhip_problem *pb, const double *uu, const double *pp, double rl,
                           double *unew, double *F, const double *p_old, const double *s_old,
                           double dt, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_RESIDUAL);
  r->mode = 2; r->x = uu; r->y = pp; r->sig = rl; r->z2 = unew; r->z = F;
  r->xq[0] = p_old; r->xq[1] = s_old; r->dt = dt;
  return submit();
}

static int e_mg_setup_jacobian(pfhip_mg *mg, const double *p, double dt, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_MG_SETUP);
  r->x = p; r->dt = dt;
  return submit();
}

static int e_mg_apply(pfhip_mg *mg, const double *b, double *x, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_MG_APPLY);
  r->x = b; r->z = x;
  return submit();
}

/* Once per Newton solve: the member's own small launches/copies, enqueued
 * behind the batched work.  Its BC staging is reused only after the
 * reductions of that solve, which synchronise the stream. */
static int e_bc_pressure(pfhip_problem *pb, double t, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_BC);
  r->dt = t;
  return submit();
}

static int e_saturation(pfhip_problem *pb, const double *p, double *out, int fcn, void *st)
{
  pfhip_ens_req *r = new_req(PFHIP_EOP_SAT);
  r->x = p; r->z = out; r->mode = fcn;
  return submit();
}

/* results reach a member only through the scheduler's synchronised copy;
 * a synchronisation point only has to launch what is queued */
static int e_synchronize(void) { return flush(); }

static const pfhip_ops ens_ops = {
  e_vec_scale, e_vec_linear_sum, e_vec_const, e_vec_dot, e_vec_max_norm,
  e_vec_scsteplength, e_vec_dq_sums, e_vec_multi_dot, e_vec_multi_axpy, e_vec_combine,
  e_kin_stop_norms, e_nl_function_eval, e_residual_jv, e_residual_step,
  e_mg_setup_jacobian, e_mg_apply, e_bc_pressure, e_saturation, e_synchronize,
  NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
};

/* ---- create / destroy --------------------------------------------------- */
int pfhip_ensemble_destroy(pfhip_ensemble *e)
{
  if (!e) return PFHIP_OK;
  for (int i = 0; i < e->n; i++) {
    ens_member *m = &e->m[i];
    for (int q = 0; q < 4; q++) pfhip_free(m->scratch[q]);
    free(m->stack);
    if (m->s) pfhip_solver_set_ops(m->s, &pfhip_direct_ops);
  }
  pfhip_free(e->d_res);
  pfhip_free(e->d_part);
  pfhip_host_free(e->h_res);
  pfhip_ens_stage_destroy(&e->st);
  pfhip_ens_mg_plan_destroy(e->mgplan);
  for (int k = 0; k < PFHIP_EOP_COUNT; k++) free(e->lists[k]);
  free(e->vec);
  free(e->outs);
  free(e->m);
  free(e);
  return PFHIP_OK;
}

int pfhip_ensemble_create(int n, pfhip_solver *const *solvers, pfhip_ensemble **out)
{
  if (n < 1 || !solvers || !out) { pfhip_set_error("ensemble: bad arguments"); return PFHIP_ERR_ARG; }
  for (int i = 0; i < n; i++) {
    const pfhip_solver_opts *o = pfhip_solver_options(solvers[i]);
    pfhip_problem *pb = pfhip_solver_problem(solvers[i]);
    if (o->krylov != 1 || !o->precond || o->use_jacobian || pb->overland ||
        o->krylov_dim > PFHIP_ENS_MAXQ - 2) {
      pfhip_set_error("ensemble member %d: needs flexible GMRES (krylov_dim <= %d) with a "
                      "preconditioner, the difference-quotient J.v and no overland flow",
                      i, PFHIP_ENS_MAXQ - 2);
      return PFHIP_ERR_ARG;
    }
    if (pb->krtab) {
      pfhip_set_error("ensemble member %d: a van Genuchten lookup table is not supported in "
                      "ensembles (members share one closure form per launch)", i);
      return PFHIP_ERR_ARG;
    }
    if (o->pc_matrix) {
      pfhip_set_error("ensemble member %d: ensembles use the PFSymmetric preconditioner "
                      "matrix", i);
      return PFHIP_ERR_ARG;
    }
  }
  pfhip_ensemble *e = (pfhip_ensemble *)calloc(1, sizeof(pfhip_ensemble));
  if (!e) { pfhip_set_error("out of host memory"); return PFHIP_ERR_NOMEM; }
  e->n = n;
  e->m = (ens_member *)calloc(n, sizeof(ens_member));
  e->vec = (pfhip_ens_req **)calloc(n, sizeof(pfhip_ens_req *));
  e->outs = (pfhip_ens_req **)calloc((long)n * ENS_QMAX, sizeof(pfhip_ens_req *));
  int rc = (e->m && e->vec && e->outs) ? PFHIP_OK : PFHIP_ERR_NOMEM;
  for (int k = 0; k < PFHIP_EOP_COUNT && !rc; k++) {
    e->lists[k] = (pfhip_ens_req **)calloc(n, sizeof(pfhip_ens_req *));
    if (!e->lists[k]) rc = PFHIP_ERR_NOMEM;
  }
  if (rc) { pfhip_set_error("out of host memory"); pfhip_ensemble_destroy(e); return rc; }
  for (int i = 0; i < n && !rc; i++) {
    ens_member *m = &e->m[i];
    m->s = solvers[i];
    m->pb = pfhip_solver_problem(m->s);
    m->mg = pfhip_solver_mg(m->s);
    for (int q = 0; q < 4 && !rc; q++) {
      rc = pfhip_dev_alloc(&m->scratch[q], m->pb->ng);
      if (!rc) rc = pfhip_zero(m->scratch[q], m->pb->ng, NULL);
    }
  }
  long blocks = 0;
  for (int i = 0; i < n && !rc; i++) blocks += pfhip_ens_vec_blocks(e->m[i].pb->ncell);
  if (!rc) rc = pfhip_dev_alloc(&e->d_part, blocks * PFHIP_ENS_NRES);
  if (!rc) rc = pfhip_dev_alloc(&e->d_res, (long)n * ENS_QMAX * PFHIP_ENS_NRES);
  if (!rc) rc = pfhip_host_alloc(&e->h_res, (long)n * ENS_QMAX * PFHIP_ENS_NRES);
  if (!rc) rc = pfhip_ens_stage_create(&e->st, ENS_STAGE);
  if (!rc) {
    pfhip_mg **mgs = (pfhip_mg **)malloc(sizeof(pfhip_mg *) * (size_t)n);
    if (!mgs) rc = PFHIP_ERR_NOMEM;
    for (int i = 0; i < n && !rc; i++) mgs[i] = e->m[i].mg;
    if (!rc) rc = pfhip_ens_mg_plan_create(mgs, n, &e->mgplan);
    free(mgs);
  }
  if (!rc) rc = pfhip_synchronize();
  if (rc) { pfhip_ensemble_destroy(e); return rc; }
  *out = e;
  return PFHIP_OK;
}

int pfhip_ensemble_stats(pfhip_ensemble *e, long *rounds, long *launches)
{
  if (rounds) *rounds = e->rounds;
  if (launches) *launches = e->launches;
  return PFHIP_OK;
}

/* ---- the scheduler ---------------------------------------------------- */
/* first frame of every coroutine: entered by the switch's return */
static void member_main(void)
{
  pfhip_ensemble *e = g_ens;
  ens_member *m = &e->m[e->cur];
  m->rc = pfhip_advance_richards(m->s, m->p, m->sat, m->t0, m->dt, m->nsteps, m->max_fail,
                                 m->newton, m->gmres, m->fevals, m->dt_used,
                                 &m->steps_done);
  /* the error message is per thread and every member shares the thread:
   * keep this member's own message before another member overwrites it */
  if (m->rc != PFHIP_OK) snprintf(m->err, sizeof(m->err), "%s", pfhip_last_error());
  int frc = flush();
  if (m->rc == PFHIP_OK && frc != PFHIP_OK) {
    m->rc = frc;
    snprintf(m->err, sizeof(m->err), "%s", pfhip_last_error());
  }
  m->live = 0;
  pfhip_ens_switch(&m->sp, e->sched_sp);   /* never resumed */
  abort();
}

/* a fresh stack whose first switch-in "returns" into member_main with the
 * alignment of a called function */
static void *coroutine_start(void *stack, long size)
{
  uintptr_t top = ((uintptr_t)stack + (uintptr_t)size) & ~(uintptr_t)15;
  void **sp = (void **)top;
  *--sp = NULL;                          /* member_main's return slot      */
  *--sp = (void *)member_main;           /* consumed by the switch's ret   */
  for (int r = 0; r < 6; r++) *--sp = NULL;   /* rbp rbx r12 r13 r14 r15   */
  return (void *)sp;
}

static void set_rc(pfhip_ens_req **l, int n, int rc)
{
  for (int i = 0; i < n; i++) l[i]->rc = rc;
}

/* Serve the queued requests of this round, position by position (a
 * member's queue runs in its order; every position is one launch per kind
 * across the members), then fetch all results with one copy.  The heavy
 * kinds (PC setup, V-cycle: their cost is set by the largest member, not by
 * how many members take part) wait while any light request is queued: the
 * members that are behind catch up through light rounds, and each heavy
 * launch then serves nearly every member at once. */
static int serve_round(pfhip_ensemble *e)
{
  int light = 0, maxp = 0, rc;
  for (int i = 0; i < e->n; i++) {
    ens_member *m = &e->m[i];
    if (m->live && m->waiting && (m->qlen > 1 || !is_heavy(m->q[m->qlen - 1].op))) light = 1;
  }
  for (int i = 0; i < e->n; i++) {
    ens_member *m = &e->m[i];
    m->nexec = 0;
    if (!m->live || !m->waiting) continue;
    m->nexec = m->qlen;
    if (light && is_heavy(m->q[m->qlen - 1].op)) m->nexec--;   /* deferred */
    if (m->nexec > maxp) maxp = m->nexec;
  }
  int slots = 0, nouts = 0;
  for (int pos = 0; pos < maxp; pos++) {
    int nv = 0;
    for (int k = 0; k < PFHIP_EOP_COUNT; k++) e->nlist[k] = 0;
    for (int i = 0; i < e->n; i++) {
      ens_member *m = &e->m[i];
      if (pos >= m->nexec) continue;
      pfhip_ens_req *r = &m->q[pos];
      if (r->op >= PFHIP_EOP_LIN && r->op <= PFHIP_EOP_STOP) {
        r->slot = slots + nv;
        e->vec[nv++] = r;
        if (r->nout) e->outs[nouts++] = r;
      } else {
        e->lists[r->op][e->nlist[r->op]++] = r;
      }
    }
    if (nv) {
      rc = pfhip_ens_exec_vec(e->vec, nv, e->d_res + (long)slots * PFHIP_ENS_NRES, e->d_part,
                              &e->st, NULL);
      set_rc(e->vec, nv, rc);
      slots += nv;
      e->launches++;
    }
    int nr = e->nlist[PFHIP_EOP_RESIDUAL];
    if (nr) {
      rc = pfhip_ens_exec_residual(e->lists[PFHIP_EOP_RESIDUAL], nr, &e->st, NULL);
      set_rc(e->lists[PFHIP_EOP_RESIDUAL], nr, rc);
      e->launches += 2;
    }
    int ns = e->nlist[PFHIP_EOP_MG_SETUP];
    if (ns) {
      rc = pfhip_ens_exec_jacobian(e->lists[PFHIP_EOP_MG_SETUP], ns, &e->st, NULL);
      long l0 = pfhip_ens_mg_plan_launches(e->mgplan);
      if (!rc) rc = pfhip_ens_exec_mg_coarse(e->mgplan, e->lists[PFHIP_EOP_MG_SETUP], ns, &e->st,
                                             NULL);
      set_rc(e->lists[PFHIP_EOP_MG_SETUP], ns, rc);
      e->launches += 2 + (pfhip_ens_mg_plan_launches(e->mgplan) - l0);
    }
    int na = e->nlist[PFHIP_EOP_MG_APPLY];
    if (na) {
      long l0 = pfhip_ens_mg_plan_launches(e->mgplan);
      rc = pfhip_ens_exec_mg_apply(e->mgplan, e->lists[PFHIP_EOP_MG_APPLY], na, &e->st, NULL);
      set_rc(e->lists[PFHIP_EOP_MG_APPLY], na, rc);
      e->launches += pfhip_ens_mg_plan_launches(e->mgplan) - l0;
    }
    int nb = e->nlist[PFHIP_EOP_BC];
    if (nb) {
      rc = pfhip_ens_exec_bc(e->lists[PFHIP_EOP_BC], nb, &e->st, NULL);
      set_rc(e->lists[PFHIP_EOP_BC], nb, rc);
      e->launches++;
    }
    int nsat = e->nlist[PFHIP_EOP_SAT];
    if (nsat) {
      rc = pfhip_ens_exec_saturation(e->lists[PFHIP_EOP_SAT], nsat, &e->st, NULL);
      set_rc(e->lists[PFHIP_EOP_SAT], nsat, rc);
      e->launches++;
    }
  }
  int frc = PFHIP_OK;
  if (nouts) {
    frc = pfhip_ens_fetch(e->h_res, e->d_res, (long)slots * PFHIP_ENS_NRES, NULL);
    if (frc == PFHIP_OK) {
      e->st.off = 0;
      for (int v = 0; v < nouts; v++) {
        pfhip_ens_req *r = e->outs[v];
        for (int k = 0; k < r->nout; k++)
          *r->out[k] = e->h_res[(long)r->slot * PFHIP_ENS_NRES + k];
      }
    }
  }
  for (int i = 0; i < e->n; i++) {
    ens_member *m = &e->m[i];
    if (!m->live || !m->waiting) continue;
    m->rc_round = frc;
    for (int p = 0; p < m->nexec && m->rc_round == PFHIP_OK; p++) m->rc_round = m->q[p].rc;
    if (m->nexec == m->qlen) {
      m->qlen = 0;
      m->waiting = 0;
    } else if (m->nexec > 0) {
      m->q[0] = m->q[m->qlen - 1];   /* the deferred heavy request */
      m->qlen = 1;
    }
  }
  e->rounds++;
  return frc;
}

int pfhip_ensemble_advance_richards(pfhip_ensemble *e, double *const *p, double *const *sat,
                                    const double *t0, const double *dt, int nsteps,
                                    int max_conv_failures, int *newton, int *gmres,
                                    int *fevals, int *steps_done)
{
  if (!e || !p || !sat || !t0 || !dt || nsteps < 0) {
    pfhip_set_error("ensemble advance: bad arguments");
    return PFHIP_ERR_ARG;
  }
  if (g_ens) { pfhip_set_error("ensemble advance: already running on this thread"); return PFHIP_ERR_ARG; }
  g_ens = e;
  e->rounds = e->launches = 0;
  int rc = PFHIP_OK;
  for (int i = 0; i < e->n; i++) {
    ens_member *m = &e->m[i];
    m->p = p[i]; m->sat = sat[i]; m->t0 = t0[i]; m->dt = dt[i];
    m->nsteps = nsteps; m->max_fail = max_conv_failures;
    m->newton = newton ? newton + (long)i * nsteps : NULL;
    m->gmres = gmres ? gmres + (long)i * nsteps : NULL;
    m->fevals = fevals ? fevals + (long)i * nsteps : NULL;
    m->dt_used = NULL;
    m->steps_done = 0;
    m->rc = PFHIP_OK;
    m->err[0] = 0;
    if (!m->stack) m->stack = malloc(ENS_STACK);
    if (!m->stack) { rc = PFHIP_ERR_NOMEM; break; }
    m->sp = coroutine_start(m->stack, ENS_STACK);
    m->waiting = 0;
    m->qlen = 0;
    m->rc_round = PFHIP_OK;
    pfhip_solver_set_ops(m->s, &ens_ops);
    m->live = 1;
  }
  if (rc) {
    for (int i = 0; i < e->n; i++) { e->m[i].live = 0; pfhip_solver_set_ops(e->m[i].s, &pfhip_direct_ops); }
    g_ens = NULL;
    pfhip_set_error("out of host memory");
    return rc;
  }
  int live = e->n;
  while (live > 0) {
    /* every live member runs until its next device request (or its end) */
    for (int i = 0; i < e->n; i++) {
      ens_member *m = &e->m[i];
      if (!m->live || m->waiting) continue;
      e->cur = i;
      pfhip_ens_switch(&e->sched_sp, m->sp);
      if (!m->live) live--;
      else m->waiting = 1;
    }
    if (live == 0) break;
    int src = serve_round(e);
    if (src && rc == PFHIP_OK) rc = src;
  }
  int first_err = PFHIP_OK, who = -1;
  for (int i = 0; i < e->n; i++) {
    ens_member *m = &e->m[i];
    pfhip_solver_set_ops(m->s, &pfhip_direct_ops);
    if (steps_done) steps_done[i] = m->steps_done;
    if (m->rc != PFHIP_OK && first_err == PFHIP_OK) { first_err = m->rc; who = i; }
  }
  g_ens = NULL;
  int sync = pfhip_synchronize();
  if (first_err != PFHIP_OK) {
    char msg[256];
    snprintf(msg, sizeof(msg), "%s", e->m[who].err[0] ? e->m[who].err : pfhip_last_error());
    pfhip_set_error("ensemble member %d: %s", who, msg);
    return first_err;
  }
  return rc ? rc : sync;
}
