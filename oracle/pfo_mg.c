/*
 * pfo_mg.c -- oracle (test infrastructure only) restatement of ParFlow's
 * in-repo semicoarsening multigrid preconditioner MGSemi:
 *   V-cycle            PL/mg_semi.c:93-400
 *   coarse operators   PL/mg_semi.c:437-760 (SetupCoarseOps)
 *   level/directions   PL/mg_semi.c:832-900
 *   restriction        PL/mg_semi_restrict.c:40-160
 *   prolongation       PL/mg_semi_prolong.c:39-160
 *   smoother           PL/rb_GS_point.c:56-380, red = (i+j+k) odd
 *                      (PL/comm_pkg.c:325-362)
 *
 * Matrices are symmetric-stored (C, E, N, U; the lower couplings are the
 * neighbour's upper ones, PL/matrix.c:258-290) or, for the FullJacobian PC
 * matrix (PL/kinsol_pc.c:215-240, PL/pf_hypre.c:300-350), full-stored
 * (C, E, N, U, W, S, L: every row keeps its own lower couplings).  All
 * arrays are ghosted (n_x+2)(n_y+2)(n_z+2) and their ghost entries stay zero.
 *
 * Full storage keeps the same cycle: operator-induced prolongation from each
 * odd point's own row, restriction by its transpose, and the coarse
 * operator as P^T A P along the coarsening direction with the off-direction
 * couplings lumped onto the diagonal (the collapse MGSemi uses), each side
 * of a row collapsed from its own couplings:
 *   Ac_U(C) = a_U(C) p2(C+1),  Ac_L(C) = a_L(C) p1(C-1),
 *   Ac_C(C) = a_0 + sum(off) + a_L(C) p2(C-1) + a_U(C) p1(C+1) - sum(Ac off),
 *   Ac_E(C) = a_E(C) + (a_E(C-1) + a_E(C+1)) / 2  (and W, N, S alike).
 * On a symmetric matrix these reduce to the MGSemi formulas above.  The
 * reference's own FullJacobian preconditioner is hypre PFMG (not vendored):
 * this hierarchy is the build's replacement for it, as for PFSymmetric.
 *
 * Parity status.  The reference's default PFMG preconditioner lives in
 * hypre (not vendored) and no reference test dumps an MGSemi iterate, so
 * the V-cycle is pinned by what the mathematics fixes: the level hierarchy
 * against the coarsening rule of PL/mg_semi.c:832-900; on a line, one
 * V-cycle from a zero guess solves A x = b to round-off (semicoarsening
 * with operator-induced weights is cyclic reduction there), which fixes
 * the coarse operators, weights, prolongation and the coarsest solve; with
 * MGSemi's palindromic red-black-red sweeps the cycle is a symmetric
 * operator, which fixes the restriction as the prolongation's transpose;
 * the cycle is a convergent solver on a 7-point Poisson operator; and the
 * end-to-end Newton solves reproduce the reference's goldens
 * (test/correct_output/default_richards.out.*, overland_tiltedV_KWE).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "pfo.h"

#define MAXL 128

struct pfo_mg {
  int L;
  int n[MAXL][3];
  int dir[MAXL];
  double *A[MAXL][7];          /* C E N U, then W S L when full */
  double *x[MAXL], *b[MAXL], *t[MAXL];
  double *p1[MAXL], *p2[MAXL];
  int pre, post, symmetric;
  int full;                    /* full (non-symmetric) storage */
};

static long lsize(const int *n) { return (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2); }

pfo_mg *pfo_mg_create_full(int nx, int ny, int nz, double dx, double dy, double dz,
                           int max_levels, int pre, int post, int symmetric, int full)
{
  pfo_mg *mg = (pfo_mg*)calloc(1, sizeof(pfo_mg));
  mg->full = full;
  int N[3] = { nx, ny, nz };
  double D[3] = { dx, dy, dz };
  int min_N = 2;          /* pfmax(2*MaxMinNX-1, 2) with MaxMinNX = 1 */
  if (max_levels <= 0 || max_levels > MAXL) max_levels = MAXL;
  mg->pre = pre;
  mg->post = post;
  mg->symmetric = symmetric;
  mg->n[0][0] = nx; mg->n[0][1] = ny; mg->n[0][2] = nz;
  int l;
  for (l = 0; l < max_levels - 1; l++)
  {
    double min_spacing = D[0] + D[1] + D[2];
    int d = -1;
    for (int q = 0; q < 3; q++)
      if ((N[q] >= min_N) && (D[q] < min_spacing))
      {
        min_spacing = D[q];
        d = q;
      }
    if (d == -1) break;
    D[d] *= 2;
    N[d] = (N[d] + 1) / 2;
    mg->dir[l] = d;
    mg->n[l + 1][0] = N[0]; mg->n[l + 1][1] = N[1]; mg->n[l + 1][2] = N[2];
  }
  mg->L = l + 1;
  for (l = 0; l < mg->L; l++)
  {
    long s = lsize(mg->n[l]);
    for (int q = 0; q < (full ? 7 : 4); q++) mg->A[l][q] = (double*)calloc(s, sizeof(double));
    mg->x[l] = (double*)calloc(s, sizeof(double));
    mg->b[l] = (double*)calloc(s, sizeof(double));
    mg->t[l] = (double*)calloc(s, sizeof(double));
    mg->p1[l] = (double*)calloc(s, sizeof(double));
    mg->p2[l] = (double*)calloc(s, sizeof(double));
  }
  return mg;
}

pfo_mg *pfo_mg_create(int nx, int ny, int nz, double dx, double dy, double dz,
                      int max_levels, int pre, int post, int symmetric)
{
  return pfo_mg_create_full(nx, ny, nz, dx, dy, dz, max_levels, pre, post, symmetric, 0);
}

void pfo_mg_destroy(pfo_mg *mg)
{
  for (int l = 0; l < mg->L; l++)
  {
    for (int q = 0; q < 7; q++) free(mg->A[l][q]);
    free(mg->x[l]); free(mg->b[l]); free(mg->t[l]);
    free(mg->p1[l]); free(mg->p2[l]);
  }
  free(mg);
}

int pfo_mg_num_levels(const pfo_mg *mg) { return mg->L; }

void pfo_mg_level_dims(const pfo_mg *mg, int l, int *dims, int *dir)
{
  dims[0] = mg->n[l][0]; dims[1] = mg->n[l][1]; dims[2] = mg->n[l][2];
  *dir = (l < mg->L - 1) ? mg->dir[l] : -1;
}

/* strides of a level */
static inline void strides(const int *n, long *s)
{
  s[0] = 1; s[1] = n[0] + 2; s[2] = (long)(n[0] + 2) * (n[1] + 2);
}
static inline long lidx(const long *s, int i, int j, int k)
{
  return (i + 1) + (j + 1) * s[1] + (k + 1) * s[2];
}

/* coupling from point ip in direction q (0,1,2) on side -1/+1; zero when
 * the neighbour lies outside the level (its ghost coefficient is zero).
 * A[4] is non-null exactly for full storage. */
static inline double cpl(double **A, const long *s, long ip, int q, int side)
{
  if (side > 0) return A[1 + q][ip];
  return A[4] ? A[4 + q][ip] : A[1 + q][ip - s[q]];
}

/* the level matrix (C, E, N, U[, W, S, L]) of level l, copied out */
void pfo_mg_level_matrix(const pfo_mg *mg, int l, double **out)
{
  long s = lsize(mg->n[l]);
  for (int q = 0; q < (mg->full ? 7 : 4); q++)
    if (out[q]) memcpy(out[q], mg->A[l][q], sizeof(double) * s);
}

void pfo_mg_setup(pfo_mg *mg, double **Asym)
{
  /* level 0 matrix = PC matrix */
  {
    long sz0 = lsize(mg->n[0]);
    memcpy(mg->A[0][0], Asym[0], sizeof(double) * sz0);
    memcpy(mg->A[0][1], Asym[2], sizeof(double) * sz0);
    memcpy(mg->A[0][2], Asym[4], sizeof(double) * sz0);
    memcpy(mg->A[0][3], Asym[6], sizeof(double) * sz0);
    if (mg->full)
    {
      memcpy(mg->A[0][4], Asym[1], sizeof(double) * sz0);
      memcpy(mg->A[0][5], Asym[3], sizeof(double) * sz0);
      memcpy(mg->A[0][6], Asym[5], sizeof(double) * sz0);
    }
  }
  for (int l = 0; l < mg->L - 1; l++)
  {
    const int *nf = mg->n[l], *nc = mg->n[l + 1];
    int d = mg->dir[l];
    long sf[3], sc[3];
    strides(nf, sf);
    strides(nc, sc);
    double **A = mg->A[l];
    double **Ac = mg->A[l + 1];
    double *p1 = mg->p1[l], *p2 = mg->p2[l];
    long lf = lsize(nf), lc = lsize(nc);
    memset(p1, 0, sizeof(double) * lf);
    memset(p2, 0, sizeof(double) * lf);
    for (int q = 0; q < (mg->full ? 7 : 4); q++) memset(Ac[q], 0, sizeof(double) * lc);
    /* off-direction order of the a3..a6 entries, PL/mg_semi.c:470-492 */
    int o1, o2;
    if (d == 0) { o1 = 1; o2 = 2; }
    else if (d == 1) { o1 = 0; o2 = 2; }
    else { o1 = 1; o2 = 0; }

    /* prolongation weights at fine (odd) points, PL/mg_semi.c:505-545 */
    for (int k = 0; k < nf[2]; k++)
      for (int j = 0; j < nf[1]; j++)
        for (int i = 0; i < nf[0]; i++)
        {
          int id[3] = { i, j, k };
          if ((id[d] & 1) == 0) continue;
          long ip = lidx(sf, i, j, k);
          double a0 = A[0][ip];
          double a1 = cpl(A, sf, ip, d, -1), a2 = cpl(A, sf, ip, d, +1);
          double a3 = cpl(A, sf, ip, o1, -1), a4 = cpl(A, sf, ip, o1, +1);
          double a5 = cpl(A, sf, ip, o2, -1), a6 = cpl(A, sf, ip, o2, +1);
          double ap0 = a0 + a3 + a4 + a5 + a6;
          if (ap0)
          {
            p1[ip] = -a1 / ap0;
            p2[ip] = -a2 / ap0;
          }
          else
          {
            p1[ip] = 0.0;
            p2[ip] = 0.0;
          }
        }

    /* coarse operator, PL/mg_semi.c:560-700 */
    for (int kc = 0; kc < nc[2]; kc++)
      for (int jc = 0; jc < nc[1]; jc++)
        for (int ic = 0; ic < nc[0]; ic++)
        {
          int fi[3] = { ic, jc, kc };
          fi[d] *= 2;
          long iA = lidx(sf, fi[0], fi[1], fi[2]);
          long iAc = lidx(sc, ic, jc, kc);
          long iA1 = iA - sf[d], iA2 = iA + sf[d];
          /* upper off-direction couplings (o1+, o2+) */
          double ac4 = A[1 + o1][iA] + 0.5 * A[1 + o1][iA1] + 0.5 * A[1 + o1][iA2];
          double ac6 = A[1 + o2][iA] + 0.5 * A[1 + o2][iA1] + 0.5 * A[1 + o2][iA2];
          Ac[1 + o1][iAc] = ac4;
          Ac[1 + o2][iAc] = ac6;
          int cid[3] = { ic, jc, kc };
          if (mg->full)
          {
            /* lower off-direction couplings (o1-, o2-) and both couplings
             * along d, each from the row's own side */
            Ac[4 + o1][iAc] = A[4 + o1][iA] + 0.5 * A[4 + o1][iA1] + 0.5 * A[4 + o1][iA2];
            Ac[4 + o2][iAc] = A[4 + o2][iA] + 0.5 * A[4 + o2][iA1] + 0.5 * A[4 + o2][iA2];
            Ac[1 + d][iAc] = (cid[d] + 1 < nc[d]) ? A[1 + d][iA] * p2[iA2] : 0.0;
            Ac[4 + d][iAc] = (cid[d] > 0) ? A[4 + d][iA] * p1[iA1] : 0.0;
          }
          /* coupling along d to the next coarse point */
          else if (cid[d] + 1 < nc[d])
          {
            long iAn = iA + 2 * sf[d];
            Ac[1 + d][iAc] = cpl(A, sf, iAn, d, -1) * p1[iA + sf[d]];
          }
          else
          {
            Ac[1 + d][iAc] = cpl(A, sf, iA, d, +1) * p2[iA + sf[d]];
          }
          double a0 = A[0][iA];
          double a1 = cpl(A, sf, iA, d, -1), a2 = cpl(A, sf, iA, d, +1);
          double a3 = cpl(A, sf, iA, o1, -1), a4 = cpl(A, sf, iA, o1, +1);
          double a5 = cpl(A, sf, iA, o2, -1), a6 = cpl(A, sf, iA, o2, +1);
          Ac[0][iAc] = a0 + a3 + a4 + a5 + a6 + a1 * p2[iA1] + a2 * p1[iA2];
        }
    /* complete the centre coefficient, PL/mg_semi.c:720-750 */
    for (int kc = 0; kc < nc[2]; kc++)
      for (int jc = 0; jc < nc[1]; jc++)
        for (int ic = 0; ic < nc[0]; ic++)
        {
          long iAc = lidx(sc, ic, jc, kc);
          double c3 = cpl(Ac, sc, iAc, o1, -1), c4 = cpl(Ac, sc, iAc, o1, +1);
          double c5 = cpl(Ac, sc, iAc, o2, -1), c6 = cpl(Ac, sc, iAc, o2, +1);
          /* coarse couplings that leave the coarse level are zero */
          int cid[3] = { ic, jc, kc };
          if (cid[o1] == 0) c3 = 0.0;
          if (cid[o1] == nc[o1] - 1) c4 = 0.0;
          if (cid[o2] == 0) c5 = 0.0;
          if (cid[o2] == nc[o2] - 1) c6 = 0.0;
          Ac[0][iAc] -= (c3 + c4 + c5 + c6);
        }
    /* zero any coupling that points out of the coarse level */
    for (int kc = 0; kc < nc[2]; kc++)
      for (int jc = 0; jc < nc[1]; jc++)
        for (int ic = 0; ic < nc[0]; ic++)
        {
          long iAc = lidx(sc, ic, jc, kc);
          int cid[3] = { ic, jc, kc };
          for (int q = 0; q < 3; q++)
          {
            if (cid[q] == nc[q] - 1) Ac[1 + q][iAc] = 0.0;
            if (mg->full && cid[q] == 0) Ac[4 + q][iAc] = 0.0;
          }
        }
  }
}

/* one RedBlackGSPoint call, PL/rb_GS_point.c:56-380 */
static void rbgs(pfo_mg *mg, int l, double *x, const double *b, int zero,
                 int max_iter)
{
  const int *n = mg->n[l];
  long s[3];
  strides(n, s);
  double **A = mg->A[l];
  int symmetric = mg->symmetric;
  int count = symmetric ? 1 : 2;
  int rb = 0;   /* 0 = red = (i+j+k) odd */
  if (zero)
  {
    count++;
    if ((count / 2) > max_iter)
    {
      memset(x, 0, sizeof(double) * lsize(n));
      if (!symmetric) return;
    }
    for (int k = 0; k < n[2]; k++)
      for (int j = 0; j < n[1]; j++)
        for (int i = 0; i < n[0]; i++)
        {
          if (((i + j + k) & 1) != 1) continue;
          long ip = lidx(s, i, j, k);
          x[ip] = b[ip] / A[0][ip];
        }
    rb = !rb;
  }
  for (; (count / 2) <= max_iter; count++)
  {
    int par = rb ? 0 : 1;
    for (int k = 0; k < n[2]; k++)
      for (int j = 0; j < n[1]; j++)
        for (int i = 0; i < n[0]; i++)
        {
          if (((i + j + k) & 1) != par) continue;
          long ip = lidx(s, i, j, k);
          double sum = 0.0;
          /* a1..a6 order W E S N L U */
          double w = (i > 0) ? cpl(A, s, ip, 0, -1) * x[ip - 1] : 0.0;
          double e = (i + 1 < n[0]) ? A[1][ip] * x[ip + 1] : 0.0;
          double so = (j > 0) ? cpl(A, s, ip, 1, -1) * x[ip - s[1]] : 0.0;
          double no = (j + 1 < n[1]) ? A[2][ip] * x[ip + s[1]] : 0.0;
          double lo = (k > 0) ? cpl(A, s, ip, 2, -1) * x[ip - s[2]] : 0.0;
          double up = (k + 1 < n[2]) ? A[3][ip] * x[ip + s[2]] : 0.0;
          sum = w + e + so + no + lo + up;
          x[ip] = (b[ip] - sum) / A[0][ip];
        }
    rb = !rb;
  }
}

/* t = b - A x on level l */
static void residual(pfo_mg *mg, int l, const double *x, const double *b,
                     double *t)
{
  const int *n = mg->n[l];
  long s[3];
  strides(n, s);
  double **A = mg->A[l];
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++)
      {
        long ip = lidx(s, i, j, k);
        double v = A[0][ip] * x[ip];
        if (i > 0) v += cpl(A, s, ip, 0, -1) * x[ip - 1];
        if (i + 1 < n[0]) v += A[1][ip] * x[ip + 1];
        if (j > 0) v += cpl(A, s, ip, 1, -1) * x[ip - s[1]];
        if (j + 1 < n[1]) v += A[2][ip] * x[ip + s[1]];
        if (k > 0) v += cpl(A, s, ip, 2, -1) * x[ip - s[2]];
        if (k + 1 < n[2]) v += A[3][ip] * x[ip + s[2]];
        t[ip] = b[ip] - v;
      }
}

/* r_c = r_f[C] + p2[C-1]*r_f[C-1] + p1[C+1]*r_f[C+1]
 * (PL/mg_semi_restrict.c:120-150) */
static void restrict_(pfo_mg *mg, int l, const double *rf, double *rc)
{
  const int *nf = mg->n[l], *nc = mg->n[l + 1];
  int d = mg->dir[l];
  long sf[3], sc[3];
  strides(nf, sf);
  strides(nc, sc);
  const double *p1 = mg->p1[l], *p2 = mg->p2[l];
  for (int kc = 0; kc < nc[2]; kc++)
    for (int jc = 0; jc < nc[1]; jc++)
      for (int ic = 0; ic < nc[0]; ic++)
      {
        int fi[3] = { ic, jc, kc };
        fi[d] *= 2;
        long iF = lidx(sf, fi[0], fi[1], fi[2]);
        long iC = lidx(sc, ic, jc, kc);
        double lo = (fi[d] > 0) ? p2[iF - sf[d]] * rf[iF - sf[d]] : 0.0;
        double hi = (fi[d] + 1 < nf[d]) ? p1[iF + sf[d]] * rf[iF + sf[d]] : 0.0;
        rc[iC] = rf[iF] + (lo + hi);
      }
}

/* e_f[C] = e_c ; e_f[F] = p1[F] e_f[F-1] + p2[F] e_f[F+1]
 * (PL/mg_semi_prolong.c:85-160) */
static void prolong(pfo_mg *mg, int l, const double *ec, double *ef)
{
  const int *nf = mg->n[l], *nc = mg->n[l + 1];
  int d = mg->dir[l];
  long sf[3], sc[3];
  strides(nf, sf);
  strides(nc, sc);
  const double *p1 = mg->p1[l], *p2 = mg->p2[l];
  for (int kc = 0; kc < nc[2]; kc++)
    for (int jc = 0; jc < nc[1]; jc++)
      for (int ic = 0; ic < nc[0]; ic++)
      {
        int fi[3] = { ic, jc, kc };
        fi[d] *= 2;
        ef[lidx(sf, fi[0], fi[1], fi[2])] = ec[lidx(sc, ic, jc, kc)];
      }
  for (int k = 0; k < nf[2]; k++)
    for (int j = 0; j < nf[1]; j++)
      for (int i = 0; i < nf[0]; i++)
      {
        int id[3] = { i, j, k };
        if ((id[d] & 1) == 0) continue;
        long ip = lidx(sf, i, j, k);
        double lo = p1[ip] * ef[ip - sf[d]];
        double hi = (id[d] + 1 < nf[d]) ? p2[ip] * ef[ip + sf[d]] : 0.0;
        ef[ip] = lo + hi;
      }
}

static void axpy_level(pfo_mg *mg, int l, const double *e, double *x)
{
  const int *n = mg->n[l];
  long s[3];
  strides(n, s);
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++)
      {
        long ip = lidx(s, i, j, k);
        x[ip] += e[ip];
      }
}

/* one V-cycle with zero initial guess (MGSemi MaxIter = 1,
 * PL/mg_semi.c:205-330); b and x are level-0 ghosted vectors */
void pfo_mg_apply(pfo_mg *mg, const double *b, double *x)
{
  int L = mg->L;
  long s0 = lsize(mg->n[0]);
  memcpy(mg->b[0], b, sizeof(double) * s0);
  memset(mg->x[0], 0, sizeof(double) * s0);
  rbgs(mg, 0, mg->x[0], mg->b[0], 1, mg->pre);
  residual(mg, 0, mg->x[0], mg->b[0], mg->t[0]);
  restrict_(mg, 0, mg->t[0], mg->b[1]);
  int l;
  for (l = 1; l <= L - 2; l++)
  {
    rbgs(mg, l, mg->x[l], mg->b[l], 1, mg->pre);
    residual(mg, l, mg->x[l], mg->b[l], mg->t[l]);
    restrict_(mg, l, mg->t[l], mg->b[l + 1]);
  }
  /* coarse solve: RedBlackGSPoint, MaxIter 1 */
  rbgs(mg, l, mg->x[l], mg->b[l], 1, 1);
  for (l = L - 2; l >= 1; l--)
  {
    prolong(mg, l, mg->x[l + 1], mg->t[l]);
    axpy_level(mg, l, mg->t[l], mg->x[l]);
    rbgs(mg, l, mg->x[l], mg->b[l], 0, mg->post);
  }
  prolong(mg, 0, mg->x[1], mg->t[0]);
  axpy_level(mg, 0, mg->t[0], mg->x[0]);
  rbgs(mg, 0, mg->x[0], mg->b[0], 0, mg->post);
  /* copy interior result */
  const int *n = mg->n[0];
  long s[3];
  strides(n, s);
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++)
      {
        long ip = lidx(s, i, j, k);
        x[ip] = mg->x[0][ip];
      }
}
