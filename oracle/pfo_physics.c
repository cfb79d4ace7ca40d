/*
 * pfo_physics.c -- oracle (test infrastructure only) restatement of the
 * Richards residual, van Genuchten closures, boundary values, kinematic
 * overland flux, Jacobian assembly and 7-point matvec of ParFlow v3.15.0.
 *
 * Every function cites the reference lines it follows.  The residual and
 * Jacobian are written in "gather" form (each interior cell sums the fluxes
 * through its own six faces) instead of the reference's scatter with
 * PlusEquals; the face flux is computed with the same operands and the same
 * operation order from both sides, so each face value is bitwise identical
 * to the one the reference scatters.
 *
 * Boundary faces: ParFlow's permeability vectors are zero in the ghost
 * layer (ctalloc, PL/vector.c:321, never updated on a single subgrid), so
 * HarmonicMean(0,K)=0 makes the interior-loop flux across a domain face and
 * the "u_old" removed by the BC patch loops exactly zero
 * (PL/nl_function_eval.c:975-2260).  Only the "u_new" term of each BC type
 * survives; that is what is computed here.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "pfo.h"

#define IDX(i, j, k) ((i) + 1 + ((j) + 1) * sy + ((k) + 1) * sz)
#define IO2(i, j)    ((i) + 1 + ((j) + 1) * sy)

/* PL/solver.h:42-47 */
static inline double hmean(double a, double b)
{
  return (a + b) ? (2.0 * a * b) / (a + b) : 0;
}
static inline double hmean_dz(double a, double b, double c, double d)
{
  return ((c * b) + (a * d)) ? (((c + d) * a * b) / ((b * c) + (a * d))) : 0;
}
static inline double upstream(double a, double b, double c, double d)
{
  return ((a - b) >= 0) ? c : d;
}
static inline double pfmaxd(double a, double b) { return (a < b) ? b : a; }

/* KS/llnlmath.c:56-62 */
static inline double rpowerr(double base, double e)
{
  if (base <= 0.0) return 0.0;
  return pow(base, e);
}

/* ---------------------------------------------------------------------- */
/* van Genuchten closures                                                  */
/* ---------------------------------------------------------------------- */

/* S(p): PL/problem_saturation.c:265-285 ; dS: :288-310 */
static inline double vg_sat(double p, double rho, double g, double alpha,
                            double n, double sres, double ssat)
{
  double m = 1.0e0 - (1.0e0 / n);
  double s_dif = ssat - sres;
  if (p >= 0.0) return s_dif + sres;
  double head = fabs(p) / (rho * g);
  return s_dif / pow(1.0 + pow((alpha * head), n), m) + sres;
}

static inline double vg_dsat(double p, double rho, double g, double alpha,
                             double n, double sres, double ssat)
{
  double m = 1.0e0 - (1.0e0 / n);
  double s_dif = ssat - sres;
  if (p >= 0.0) return 0.0;
  double head = fabs(p) / (rho * g);
  return (m * n * alpha * pow(alpha * head, (n - 1))) * s_dif
         / (pow(1.0 + pow(alpha * head, n), m + 1));
}

/* ---------------------------------------------------------------------- */
/* van Genuchten kr lookup table                                           */
/* ---------------------------------------------------------------------- */

/* VanGComputeTable (PL/problem_phase_rel_perm.c:116-262): kr and its
 * derivative formula at points+1 evenly spaced heads x_i = i |min_head| /
 * (points-1), the Fritsch-Carlson monotone slopes d of both, and the
 * linear slopes.  tab holds 7 arrays of points+1 values:
 * x | a | d | a_der | d_der | slope | slope_der. */
void pfo_vg_table_build(int method, int points, double min_head, double alpha,
                        double n, double *tab)
{
  const int N = points;
  double *x = tab, *a = tab + (N + 1), *d = tab + 2 * (N + 1);
  double *a_der = tab + 3 * (N + 1), *d_der = tab + 4 * (N + 1);
  double *slope = tab + 5 * (N + 1), *slope_der = tab + 6 * (N + 1);
  double interval = min_head / (double)(N - 1);
  double ainterval = fabs(interval);
  double m = 1.0e0 - (1.0e0 / n);
  for (int i = 0; i <= N; i++)
  {
    x[i] = fabs(i * interval);
    double opahn = 1.0 + pow(alpha * x[i], n);
    double ahnm1 = pow(alpha * x[i], n - 1);
    a[i] = pow(1.0 - ahnm1 / (pow(opahn, m)), 2) / pow(opahn, (m / 2));
    double coeff = 1.0 - ahnm1 * pow(opahn, -m);
    a_der[i] = 2.0 * (coeff / (pow(opahn, (m / 2))))
               * ((n - 1) * pow(alpha * x[i], n - 2) * alpha * pow(opahn, -m)
                  - ahnm1 * m * pow(opahn, -(m + 1)) * n * alpha * ahnm1)
               + pow(coeff, 2) * (m / 2) * pow(opahn, (-(m + 2) / 2))
               * n * alpha * ahnm1;
    /* 1 < n < 2: the derivative is infinite at head 0 (:181-184) */
    if ((n < 2) && (i == 0)) a_der[i] = 0;
    slope[i] = 0.0;
    slope_der[i] = 0.0;
  }
  if (method == 1)
    for (int i = 0; i < N; i++)
    {
      slope[i] = (a[i + 1] - a[i]) / ainterval;
      slope_der[i] = (a_der[i + 1] - a_der[i]) / ainterval;
    }
  /* monotone cubic Hermite slopes, Fritsch & Carlson (1980) (:201-259) */
  double *del = (double*)malloc(sizeof(double) * (N + 1));
  double *del_der = (double*)malloc(sizeof(double) * (N + 1));
  for (int i = 0; i < N; i++)
  {
    double h = x[i + 1] - x[i];
    del[i] = (a[i + 1] - a[i]) / h;
    del_der[i] = (a_der[i + 1] - a_der[i]) / h;
  }
  d[0] = del[0];
  d[N] = del[N - 1];
  d_der[0] = del_der[0];
  d_der[N] = del_der[N - 1];
  for (int i = 1; i < N; i++)
  {
    d[i] = (del[i - 1] + del[i]) / 2;
    d_der[i] = (del_der[i - 1] + del_der[i]) / 2;
  }
  for (int i = 0; i < N; i++)
  {
    if (del[i] == 0.0)
    {
      d[i] = 0;
      d[i + 1] = 0;
    }
    else
    {
      double al = d[i] / del[i], be = d[i + 1] / del[i];
      double magn = pow(al, 2) + pow(be, 2);
      if (magn > 9.0)
      {
        d[i] = 3 * al * del[i] / magn;
        d[i + 1] = 3 * be * del[i] / magn;
      }
    }
    if (del_der[i] == 0.0)
    {
      d_der[i] = 0;
      d_der[i + 1] = 0;
    }
    else
    {
      double al = d_der[i] / del_der[i], be = d_der[i + 1] / del_der[i];
      double magn = pow(al, 2) + pow(be, 2);
      if (magn > 9.0)
      {
        d_der[i] = 3 * al * del_der[i] / magn;
        d_der[i + 1] = 3 * be * del_der[i] / magn;
      }
    }
  }
  free(del);
  free(del_der);
}

/* VanGLookupSpline (PL/problem_phase_rel_perm.c:265-354) and the linear
 * lookup of PhaseRelPerm (:718-760, :846-893); head >= 0, fcn 0 kr,
 * 1 the derivative formula's value.  0 beyond |min_head|. */
double pfo_vg_table_lookup(const double *tab, int method, int points,
                           double min_head, double head, int fcn)
{
  const int N = points;
  const double *x = tab, *a = tab + (N + 1), *d = tab + 2 * (N + 1);
  const double *a_der = tab + 3 * (N + 1), *d_der = tab + 4 * (N + 1);
  const double *slope = tab + 5 * (N + 1), *slope_der = tab + 6 * (N + 1);
  double interval = fabs(min_head / (double)(N - 1));
  if (head >= fabs(min_head)) return 0.0;
  int pt = (int)floor(head / interval);
  if (method == 1)
  {
    if (fcn == PFO_CALCFCN) return a[pt] + slope[pt] * (head - x[pt]);
    return a_der[pt] + slope_der[pt] * (head - x[pt]);
  }
  if (pt > N + 1) pt = N;
  double x0 = x[pt], h = x[pt + 1] - x0;
  double t = (head - x0) / h;
  const double *A = (fcn == PFO_CALCFCN) ? a : a_der;
  const double *D = (fcn == PFO_CALCFCN) ? d : d_der;
  return (2.0 * pow(t, 3) - 3.0 * pow(t, 2) + 1.0) * A[pt]
         + (pow(t, 3) - 2.0 * pow(t, 2) + t) * h * D[pt]
         + (-2.0 * pow(t, 3) + 3.0 * pow(t, 2)) * A[pt + 1]
         + (pow(t, 3) - pow(t, 2)) * h * D[pt + 1];
}

/* kr(p): PL/problem_phase_rel_perm.c:767-792, or the problem's table */
static inline double vg_kr(const pfo_problem *pb, double p, double rho, double g,
                           double alpha, double n)
{
  if (p >= 0.0) return 1.0;
  double head = fabs(p) / (rho * g);
  if (pb->vgt)
    return pfo_vg_table_lookup(pb->vgt, pb->vgt_method, pb->vgt_points,
                               pb->vgt_min_head, head, PFO_CALCFCN);
  double m = 1.0e0 - (1.0e0 / n);
  double opahn = 1.0 + pow(alpha * head, n);
  double ahnm1 = pow(alpha * head, n - 1);
  return pow(1.0 - ahnm1 / (pow(opahn, m)), 2) / pow(opahn, (m / 2));
}

/* dkr/dp: PL/problem_phase_rel_perm.c:895-912, or the problem's table */
static inline double vg_dkr(const pfo_problem *pb, double p, double rho, double g,
                            double alpha, double n)
{
  if (p >= 0.0) return 0.0;
  double head = fabs(p) / (rho * g);
  if (pb->vgt)
    return pfo_vg_table_lookup(pb->vgt, pb->vgt_method, pb->vgt_points,
                               pb->vgt_min_head, head, PFO_CALCDER);
  double m = 1.0e0 - (1.0e0 / n);
  double opahn = 1.0 + pow(alpha * head, n);
  double ahnm1 = pow(alpha * head, n - 1);
  double coeff = 1.0 - ahnm1 * pow(opahn, -m);
  return 2.0 * (coeff / (pow(opahn, (m / 2))))
         * ((n - 1) * pow(alpha * head, n - 2) * alpha
            * pow(opahn, -m)
            - ahnm1 * m * pow(opahn, -(m + 1)) * n * alpha * ahnm1)
         + pow(coeff, 2) * (m / 2) * pow(opahn, (-(m + 2) / 2))
         * n * alpha * ahnm1;
}

void pfo_saturation(const pfo_problem *pb, const double *p, double *out,
                    int fcn)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        out[ip] = (fcn == PFO_CALCFCN)
                  ? vg_sat(p[ip], pb->rho, pb->gravity, pb->alpha[ip],
                           pb->nvg[ip], pb->sres[ip], pb->ssat[ip])
                  : vg_dsat(p[ip], pb->rho, pb->gravity, pb->alpha[ip],
                            pb->nvg[ip], pb->sres[ip], pb->ssat[ip]);
      }
}

void pfo_relperm(const pfo_problem *pb, const double *p, double *out, int fcn)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        out[ip] = (fcn == PFO_CALCFCN)
                  ? vg_kr(pb, p[ip], pb->rho, pb->gravity, pb->alpha[ip],
                          pb->nvg[ip])
                  : vg_dkr(pb, p[ip], pb->rho, pb->gravity, pb->alpha[ip],
                           pb->nvg[ip]);
      }
}

/* ---------------------------------------------------------------------- */
/* Boundary values                                                         */
/* ---------------------------------------------------------------------- */

/* PL/time_cycle_data.c:152-197 (TIME_CYCLE_SUBDIVISIONS = 1000) */
int pfo_interval_number(const pfo_problem *pb, const pfo_bc_spec *bc,
                        double time)
{
  int cycle_length = 0;
  for (int i = 0; i < bc->n_intervals; i++) cycle_length += bc->lengths[i];
  double bu = pb->base_unit;
  if ((bc->repeat < 0) ||
      (time < ((double)(bc->repeat * cycle_length * bu))))
  {
    int dt_ = (int)lround(time / (bu / 1000)) / 1000;
    int ds_ = (int)lround(pb->start_time / (bu / 1000)) / 1000;
    int completed = (dt_ - ds_ - 1) % cycle_length;
    int interval = 0;
    int total = bc->lengths[0];
    while (completed >= total)
    {
      interval++;
      total += bc->lengths[interval];
    }
    return interval;
  }
  return bc->n_intervals - 1;
}

static int face_ncells(const pfo_problem *pb, int f)
{
  if (f < 2) return pb->ny * pb->nz;
  if (f < 4) return pb->nx * pb->nz;
  return pb->nx * pb->ny;
}

/* BCPressure: PL/problem_bc_pressure.c:143-290 (DirEquilRefPatch with a
 * constant density: the Newton loop of :249-272 converges in one update to
 * p = p_ref - rho*g*(z - z_ref)); FluxConst :582-612; OverlandKinematic
 * :1070-1098.  z is the cell centre for side faces and the face elevation
 * for the z faces (:229-234). */
void pfo_bc_pressure(pfo_problem *pb, double time)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  for (int f = 0; f < 6; f++)
  {
    const pfo_bc_spec *bc = &pb->bc[f];
    int n = face_ncells(pb, f);
    double *v = pb->bc_val[f];
    if (!v) continue;
    int iv = pfo_interval_number(pb, bc, time);
    double val = bc->values[iv];
    if (bc->type != PFO_BC_DIR_EQUIL)
    {
      for (int c = 0; c < n; c++) v[c] = val;
      continue;
    }
    double ref_den = pb->rho;
    double gneg = -pb->gravity;
    for (int c = 0; c < n; c++)
    {
      int k;
      double z;
      if (f < 2)      k = c / ny;
      else if (f < 4) k = c / nx;
      else            k = (f == 4) ? 0 : nz - 1;
      z = pb->z0 + (k + 0.5) * pb->dz;
      if (f == 4) z = z - 0.5 * pb->dz;
      if (f == 5) z = z + 0.5 * pb->dz;
      /* Newton iteration of PL/problem_bc_pressure.c:249-272 */
      double pv = val, fcn_val, resid = 1.0;
      int it = -1;
      while ((resid > 1.0E-6) && (it < 10))
      {
        if (it > -1)
        {
          double dtmp = 1.0 - 0.5 * 0.0 * gneg * (z - bc->z_ref);
          pv = pv - fcn_val / dtmp;
        }
        else
          pv = val;
        fcn_val = pv - val - 0.5 * (pb->rho + ref_den) * gneg
                  * (z - bc->z_ref);
        resid = fabs(fcn_val);
        it++;
      }
      v[c] = pv;
    }
  }
}

/* HydroStaticPatch initial condition (PL/problem_ic_phase_pressure.c:440-673
 * with constant density): p = value - rho*g*(z_centre - z_ref). */
void pfo_hydrostatic_ic(const pfo_problem *pb, double value, double z_ref,
                        double *p)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  for (int k = 0; k < nz; k++)
  {
    double z = pb->z0 + (k + 0.5) * pb->dz;
    double fcn_val, pv = value, resid = 1.0;
    int it = -1;
    while ((resid > 1.0E-6) && (it < 10))
    {
      if (it > -1) pv = pv - fcn_val;
      else pv = value;
      fcn_val = pv - value + 0.5 * (pb->rho + pb->rho) * pb->gravity
                * (z - z_ref);
      resid = fabs(fcn_val);
      it++;
    }
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        p[IDX(i, j, k)] = pv;
  }
}

/* ---------------------------------------------------------------------- */
/* Kinematic overland flow                                                  */
/* ---------------------------------------------------------------------- */

/* Face fluxes qx (east face of column io), qy (north face of column io) for
 * a box domain whose whole top is one OverlandKinematic patch:
 * PL/overlandflow_eval_Kin.c:120-290.  Columns outside the domain have
 * top index -1, so the "lower boundary" fixes (:226-270) set the west/south
 * outflow faces and an inflow from outside carries zero depth (the ghost
 * pressure is 0, so max(p,0)=0).  qx/qy are (nx+2)(ny+2), zeroed here. */
void pfo_overland_kin(const pfo_problem *pb, const double *p, double *qx,
                      double *qy)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const double eps = pb->ov_eps;
  memset(qx, 0, sizeof(double) * sz);
  memset(qy, 0, sizeof(double) * sz);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
    {
      int io = IO2(i, j);
      int ip = IDX(i, j, nz - 1);
      double Sf_x = pb->slope_x[io], Sf_y = pb->slope_y[io];
      double Sf_mag = rpowerr(Sf_x * Sf_x + Sf_y * Sf_y, 0.5);
      if (Sf_mag < eps) Sf_mag = eps;
      double PP_ip = p[ip];
      double PP_ipp1 = (i + 1 < nx) ? p[IDX(i + 1, j, nz - 1)] : 0.0;
      double PP_ippsy = (j + 1 < ny) ? p[IDX(i, j + 1, nz - 1)] : 0.0;
      double Press_x = upstream(-Sf_x, 0.0, pfmaxd(PP_ip, 0.0),
                                pfmaxd(PP_ipp1, 0.0));
      double Press_y = upstream(-Sf_y, 0.0, pfmaxd(PP_ip, 0.0),
                                pfmaxd(PP_ippsy, 0.0));
      qx[io] = -(Sf_x / (rpowerr(fabs(Sf_mag), 0.5) * pb->mann[io]))
               * rpowerr(Press_x, (5.0 / 3.0));
      qy[io] = -(Sf_y / (rpowerr(fabs(Sf_mag), 0.5) * pb->mann[io]))
               * rpowerr(Press_y, (5.0 / 3.0));
      if (i == 0 && Sf_x > 0.0)
      {
        double Pr = pfmaxd(PP_ip, 0.0);
        qx[io - 1] = -(Sf_x / (rpowerr(fabs(Sf_mag), 0.5) * pb->mann[io]))
                     * rpowerr(Pr, (5.0 / 3.0));
      }
      if (j == 0 && Sf_y > 0.0)
      {
        double Pr = pfmaxd(PP_ip, 0.0);
        qy[io - sy] = -(Sf_y / (rpowerr(fabs(Sf_mag), 0.5) * pb->mann[io]))
                      * rpowerr(Pr, (5.0 / 3.0));
      }
    }
}

/* CALCDER branch: PL/overlandflow_eval_Kin.c:293-460.  Arrays zeroed here. */
void pfo_overland_kin_der(const pfo_problem *pb, const double *p,
                          double *ke, double *kw, double *kn, double *ks)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const double eps = pb->ov_eps;
  memset(ke, 0, sizeof(double) * sz);
  memset(kw, 0, sizeof(double) * sz);
  memset(kn, 0, sizeof(double) * sz);
  memset(ks, 0, sizeof(double) * sz);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
    {
      int io = IO2(i, j);
      int ip = IDX(i, j, nz - 1);
      double Sf_x = pb->slope_x[io], Sf_y = pb->slope_y[io];
      double Sf_mag = rpowerr(Sf_x * Sf_x + Sf_y * Sf_y, 0.5);
      if (Sf_mag < eps) Sf_mag = eps;
      double PP_ipp1 = (i + 1 < nx) ? p[IDX(i + 1, j, nz - 1)] : 0.0;
      double PP_ippsy = (j + 1 < ny) ? p[IDX(i, j + 1, nz - 1)] : 0.0;
      double Press_x = upstream(-Sf_x, 0.0, pfmaxd(p[ip], 0.0),
                                pfmaxd(PP_ipp1, 0.0));
      double Press_y = upstream(-Sf_y, 0.0, pfmaxd(p[ip], 0.0),
                                pfmaxd(PP_ippsy, 0.0));
      double qx_t = -(5.0 / 3.0) * (Sf_x / (rpowerr(fabs(Sf_mag), 0.5)
                                            * pb->mann[io]))
                    * rpowerr(Press_x, (2.0 / 3.0));
      double qy_t = -(5.0 / 3.0) * (Sf_y / (rpowerr(fabs(Sf_mag), 0.5)
                                            * pb->mann[io]))
                    * rpowerr(Press_y, (2.0 / 3.0));
      ke[io] = pfmaxd(qx_t, 0);
      kw[io + 1] = -pfmaxd(-qx_t, 0);
      kn[io] = pfmaxd(qy_t, 0);
      ks[io + sy] = -pfmaxd(-qy_t, 0);
      if (i == 0 && Sf_x > 0.0)
      {
        double Pr = pfmaxd(p[ip], 0.0);
        double q = -(5.0 / 3.0) * (Sf_x / (rpowerr(fabs(Sf_mag), 0.5)
                                           * pb->mann[io]))
                   * rpowerr(Pr, (2.0 / 3.0));
        kw[io] = -pfmaxd(-q, 0);
        ke[io - 1] = pfmaxd(q, 0);
      }
      if (j == 0 && Sf_y > 0.0)
      {
        double Pr = pfmaxd(p[ip], 0.0);
        double q = -(5.0 / 3.0) * (Sf_y / (rpowerr(fabs(Sf_mag), 0.5)
                                           * pb->mann[io]))
                   * rpowerr(Pr, (2.0 / 3.0));
        ks[io] = -pfmaxd(-q, 0);
        kn[io - sy] = pfmaxd(q, 0);
      }
    }
}

/* ---------------------------------------------------------------------- */
/* Residual                                                                */
/* ---------------------------------------------------------------------- */

/* face index helpers: value arrays per face */
static inline int fidx(const pfo_problem *pb, int f, int i, int j, int k)
{
  if (f < 2) return j + k * pb->ny;
  if (f < 4) return i + k * pb->nx;
  return i + j * pb->nx;
}

/* DirichletBC "u_new" of face fc of boundary cell ip with BC value `value`
 * (PL/nl_function_eval.c:1018-1232); the ghost kr*rho is kr at the
 * Dirichlet value inserted at :574-610.  Used for F (:1225) and for the
 * face velocity u_new / h (:1230). */
static double dir_u_new(const pfo_problem *pb, int fc, int ip, double value,
                        double pc, double krc)
{
  const double dx = pb->dx, dy = pb->dy, dz = pb->dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double rho = pb->rho, g = pb->gravity, mu = pb->mu;
  const double z_dir_g = 1.0;
  double h = (fc < 2) ? ffx : (fc < 4) ? ffy : ffz;
  double kg = vg_kr(pb, value, rho, g, pb->alpha[ip], pb->nvg[ip]) * rho;
  double u_new;
  if (fc == 0)
  {
    double diff = value - pc;
    u_new = upstream(value, pc, kg, krc);
    u_new = u_new * h * (pb->kx[ip] / mu) * 2.0 * (diff / dx);
  }
  else if (fc == 1)
  {
    double diff = pc - value;
    u_new = upstream(pc, value, krc, kg);
    u_new = u_new * h * (pb->kx[ip] / mu) * 2.0 * (diff / dx);
  }
  else if (fc == 2)
  {
    double diff = value - pc;
    u_new = upstream(value, pc, kg, krc);
    u_new = u_new * h * (pb->ky[ip] / mu) * 2.0 * (diff / dy);
  }
  else if (fc == 3)
  {
    double diff = pc - value;
    u_new = upstream(pc, value, krc, kg);
    u_new = u_new * h * (pb->ky[ip] / mu) * 2.0 * (diff / dy);
  }
  else if (fc == 4)
  {
    double sep = dz * 1.0;
    double lc = value / sep - 0.25 * rho * g;
    double uc = pc / sep + 0.25 * rho * g;
    double diff = lc - uc;
    u_new = upstream(lc, uc, kg, krc);
    u_new = u_new * h * (pb->kz[ip] / mu) * 2.0 * diff;
  }
  else
  {
    double sep = dz * 1.0;
    double lc = (pc / sep) - 0.25 * rho * g * z_dir_g;
    double uc = (value / sep) + 0.25 * rho * g * z_dir_g;
    double diff = lc - uc;
    u_new = upstream(lc, uc, krc, kg);
    u_new = u_new * h * (pb->kz[ip] / mu) * 2.0 * diff;
  }
  return u_new;
}

/* NlFunctionEval: PL/nl_function_eval.c:167-2323, gather form. */
void pfo_nl_function_eval(const pfo_problem *pb, const double *p,
                          const double *p_old, const double *s_old,
                          double dt, double *F)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const long ng = (long)sz * (nz + 2);
  const double dx = pb->dx, dy = pb->dy, dz = pb->dz;
  const double vol = dx * dy * dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double rho = pb->rho, g = pb->gravity, mu = pb->mu;
  /* x_dir_g_c = gravity*cos(atan(0)) for zero subsurface slopes,
   * PL/nl_function_eval.c:690-697 */
  const double gc = 0.5 * (g * cos(atan(0.0)) + g * cos(atan(0.0)));
  const double z_dir_g = 1.0;

  double *krho = (double*)malloc(sizeof(double) * ng);
  double *qx = NULL, *qy = NULL;
  int overland = (pb->bc[5].type == PFO_BC_OVERLAND);
  if (overland)
  {
    qx = (double*)malloc(sizeof(double) * sz);
    qy = (double*)malloc(sizeof(double) * sz);
    pfo_overland_kin(pb, p, qx, qy);
  }

  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        krho[ip] = vg_kr(pb, p[ip], rho, g, pb->alpha[ip], pb->nvg[ip]) * rho;
      }

  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        double S = vg_sat(p[ip], rho, g, pb->alpha[ip], pb->nvg[ip],
                          pb->sres[ip], pb->ssat[ip]);
        double so = s_old[ip];
        /* accumulation :394-405, compressible storage :462-472 */
        double f = (S * rho - so * rho) * pb->poro[ip] * vol;
        f += pb->ss[ip] * vol * (p[ip] * S * rho - p_old[ip] * so * rho);
        /* sources :530-540 */
        if (pb->src) f -= vol * dt * (pb->src[ip]);

        /* interior faces, :696-854 */
        double flux_own = 0.0;   /* u_right + u_front + u_upper */
        double ur = 0.0, uf = 0.0, uu = 0.0;
        if (i + 1 < nx)
        {
          int b = ip + 1;
          double diff = p[ip] - p[b];
          double updir = (diff / dx) * gc;
          double up = upstream(updir, 0.0, krho[ip], krho[b]);
          ur = ffx * hmean(pb->kx[ip], pb->kx[b]) * (diff / dx) * gc * up / mu;
        }
        if (j + 1 < ny)
        {
          int b = ip + sy;
          double diff = p[ip] - p[b];
          double updir = (diff / dy) * gc;
          double up = upstream(updir, 0.0, krho[ip], krho[b]);
          uf = ffy * hmean(pb->ky[ip], pb->ky[b]) * (diff / dy) * gc * up / mu;
        }
        if (k + 1 < nz)
        {
          int b = ip + sz;
          double sep = dz;
          double lc = p[ip] / sep - 0.5 * rho * g * z_dir_g;
          double uc = p[b] / sep + 0.5 * rho * g * z_dir_g;
          double diff = lc - uc;
          uu = ffz * hmean_dz(pb->kz[ip], pb->kz[b], 1.0, 1.0) * diff
               * upstream(lc, uc, krho[ip], krho[b]) / mu;
        }
        flux_own = ur + uf + uu;
        f += dt * flux_own;
        if (i > 0)
        {
          int a = ip - 1;
          double diff = p[a] - p[ip];
          double updir = (diff / dx) * gc;
          double up = upstream(updir, 0.0, krho[a], krho[ip]);
          double u = ffx * hmean(pb->kx[a], pb->kx[ip]) * (diff / dx) * gc * up / mu;
          f += -dt * u;
        }
        if (j > 0)
        {
          int a = ip - sy;
          double diff = p[a] - p[ip];
          double updir = (diff / dy) * gc;
          double up = upstream(updir, 0.0, krho[a], krho[ip]);
          double u = ffy * hmean(pb->ky[a], pb->ky[ip]) * (diff / dy) * gc * up / mu;
          f += -dt * u;
        }
        if (k > 0)
        {
          int a = ip - sz;
          double sep = dz;
          double lc = p[a] / sep - 0.5 * rho * g * z_dir_g;
          double uc = p[ip] / sep + 0.5 * rho * g * z_dir_g;
          double diff = lc - uc;
          double u = ffz * hmean_dz(pb->kz[a], pb->kz[ip], 1.0, 1.0) * diff
                     * upstream(lc, uc, krho[a], krho[ip]) / mu;
          f += -dt * u;
        }

        /* boundary faces :975-2260 */
        for (int fc = 0; fc < 6; fc++)
        {
          int on = (fc == 0 && i == 0) || (fc == 1 && i == nx - 1) ||
                   (fc == 2 && j == 0) || (fc == 3 && j == ny - 1) ||
                   (fc == 4 && k == 0) || (fc == 5 && k == nz - 1);
          if (!on) continue;
          int type = pb->bc[fc].type;
          double value = pb->bc_val[fc][fidx(pb, fc, i, j, k)];
          int dir = (fc & 1) ? 1 : -1;
          double h = (fc < 2) ? ffx : (fc < 4) ? ffy : ffz;
          if (type == PFO_BC_FLUX)
          {
            /* FluxBC :1413-1422 */
            double u_new = h * value;
            f += dt * dir * u_new;
          }
          else if (type == PFO_BC_DIR_EQUIL)
          {
            double u_new = dir_u_new(pb, fc, ip, value, p[ip], krho[ip]);
            f += dt * dir * u_new;
          }
          else if (type == PFO_BC_OVERLAND && fc == 5)
          {
            /* OverlandKinematicBC :1786-2010 */
            int io = IO2(i, j);
            double ke = qx[io], kw = qx[io - 1];
            double kn = qy[io], ks = qy[io - sy];
            double q_overlnd = vol * (pfmaxd(p[ip], 0.0)
                                      - pfmaxd(p_old[ip], 0.0)) / dz
                               + dt * vol * ((ke - kw) / dx + (kn - ks) / dy)
                               / dz;
            f += q_overlnd;
            double u_new = h * value;
            f += dt * dir * u_new;
          }
        }
        F[ip] = f;
      }

  free(krho);
  if (qx) free(qx);
  if (qy) free(qy);
}

/* ---------------------------------------------------------------------- */
/* Face velocities                                                         */
/* ---------------------------------------------------------------------- */

/* The x/y/z face velocities NlFunctionEval writes beside F
 * (PL/nl_function_eval.c:710-712, :846-848 interior; :1230 DirichletBC,
 * :1432 FluxBC, :2023 OverlandKinematicBC), stored face-centred without
 * ghosts, x fastest:
 *   vx[i + (nx+1)*(j + ny*k)], face i between cells i-1 and i (0..nx)
 *   vy[i + nx*(j + (ny+1)*k)], face j (0..ny)
 *   vz[i + nx*(j + ny*k)],     face k (0..nz)
 * the layout of the reference's velx/vely/velz PFB outputs.  Interior
 * faces: the face flux u of the residual divided by the face area
 * (ffx * z_mult * del_y_slope = ffx with z_mult = 1 and zero slopes).
 * Domain faces: the value the BC patch loop leaves (u_new / h for
 * Dirichlet, the flux value for FluxConst, (h value + q_overlnd / dt) / h
 * on a kinematic overland top).  p_old and dt enter only through q_overlnd. */
void pfo_face_velocity(const pfo_problem *pb, const double *p,
                       const double *p_old, double dt, double *vx,
                       double *vy, double *vz)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const long ng = (long)sz * (nz + 2);
  const double dx = pb->dx, dy = pb->dy, dz = pb->dz;
  const double vol = dx * dy * dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double rho = pb->rho, g = pb->gravity, mu = pb->mu;
  const double gc = 0.5 * (g * cos(atan(0.0)) + g * cos(atan(0.0)));
  const double z_dir_g = 1.0;
  double *krho = (double*)calloc(ng, sizeof(double));
  double *qx = NULL, *qy = NULL;
  if (pb->bc[5].type == PFO_BC_OVERLAND)
  {
    qx = (double*)malloc(sizeof(double) * sz);
    qy = (double*)malloc(sizeof(double) * sz);
    pfo_overland_kin(pb, p, qx, qy);
  }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        krho[ip] = vg_kr(pb, p[ip], rho, g, pb->alpha[ip], pb->nvg[ip]) * rho;
      }

  /* value the BC patch loop of face fc leaves for boundary cell (i,j,k) */
  #define BC_VEL(fc, i, j, k, out)                                          \
  do {                                                                      \
    int ip_ = IDX(i, j, k);                                                 \
    double value_ = pb->bc_val[fc][fidx(pb, fc, i, j, k)];                  \
    double h_ = ((fc) < 2) ? ffx : ((fc) < 4) ? ffy : ffz;                  \
    int type_ = pb->bc[fc].type;                                            \
    if (type_ == PFO_BC_FLUX) (out) = value_;                               \
    else if (type_ == PFO_BC_DIR_EQUIL)                                     \
      (out) = dir_u_new(pb, fc, ip_, value_, p[ip_], krho[ip_]) / h_;       \
    else {                                                                  \
      int io_ = IO2(i, j);                                                  \
      double q_ = vol * (pfmaxd(p[ip_], 0.0) - pfmaxd(p_old[ip_], 0.0)) / dz \
                  + dt * vol * ((qx[io_] - qx[io_ - 1]) / dx                \
                                + (qy[io_] - qy[io_ - sy]) / dy) / dz;      \
      double u_ = h_ * value_;                                              \
      u_ += q_ / dt;                                                        \
      (out) = u_ / h_;                                                      \
    }                                                                       \
  } while (0)

  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++)
      {
        double v;
        if (i == 0) BC_VEL(0, 0, j, k, v);
        else if (i == nx) BC_VEL(1, nx - 1, j, k, v);
        else
        {
          int a = IDX(i - 1, j, k), b = IDX(i, j, k);
          double diff = p[a] - p[b];
          double updir = (diff / dx) * gc;
          double up = upstream(updir, 0.0, krho[a], krho[b]);
          double u = ffx * hmean(pb->kx[a], pb->kx[b]) * (diff / dx) * gc * up / mu;
          v = u / (ffx * 1.0 * 1.0);
        }
        vx[i + (long)(nx + 1) * (j + (long)ny * k)] = v;
      }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++)
      {
        double v;
        if (j == 0) BC_VEL(2, i, 0, k, v);
        else if (j == ny) BC_VEL(3, i, ny - 1, k, v);
        else
        {
          int a = IDX(i, j - 1, k), b = IDX(i, j, k);
          double diff = p[a] - p[b];
          double updir = (diff / dy) * gc;
          double up = upstream(updir, 0.0, krho[a], krho[b]);
          double u = ffy * hmean(pb->ky[a], pb->ky[b]) * (diff / dy) * gc * up / mu;
          v = u / (ffy * 1.0 * 1.0);
        }
        vy[i + (long)nx * (j + (long)(ny + 1) * k)] = v;
      }
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        double v;
        if (k == 0) BC_VEL(4, i, j, 0, v);
        else if (k == nz) BC_VEL(5, i, j, nz - 1, v);
        else
        {
          int a = IDX(i, j, k - 1), b = IDX(i, j, k);
          double sep = dz;
          double lc = p[a] / sep - 0.5 * rho * g * z_dir_g;
          double uc = p[b] / sep + 0.5 * rho * g * z_dir_g;
          double diff = lc - uc;
          double u = ffz * hmean_dz(pb->kz[a], pb->kz[b], 1.0, 1.0) * diff
                     * upstream(lc, uc, krho[a], krho[b]) / mu;
          v = u / (ffz * 1.0 * 1.0);
        }
        vz[i + (long)nx * (j + (long)ny * k)] = v;
      }
  #undef BC_VEL
  free(krho);
  if (qx) free(qx);
  if (qy) free(qy);
}

/* Water-balance terms of one step (volumes over the step; inflow positive,
 * runoff positive leaving): [0] domain faces other than an overland top,
 * from the face velocities above; [1] the overland top's flux BC;
 * [2] kinematic runoff through the domain edges, from the surface fluxes
 * of pfo_overland_kin; [3] sum src vol dt (PL/evaptranssum.c:86-92).
 * The budget of test/python/water_balance_x.py:546-620. */
void pfo_water_balance_fluxes(const pfo_problem *pb, const double *p,
                              const double *p_old, double dt, double *out)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const double dx = pb->dx, dy = pb->dy, dz = pb->dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const int overland = (pb->bc[5].type == PFO_BC_OVERLAND);
  double *vx = (double*)malloc(sizeof(double) * (nx + 1) * ny * nz);
  double *vy = (double*)malloc(sizeof(double) * nx * (ny + 1) * nz);
  double *vz = (double*)malloc(sizeof(double) * nx * ny * (nz + 1));
  pfo_face_velocity(pb, p, p_old, dt, vx, vy, vz);
  double bin = 0.0, rain = 0.0, run = 0.0, src = 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
    {
      bin += dt * ffx * vx[0 + (nx + 1) * (j + ny * k)];
      bin -= dt * ffx * vx[nx + (nx + 1) * (j + ny * k)];
    }
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nx; i++)
    {
      bin += dt * ffy * vy[i + nx * (0 + (ny + 1) * k)];
      bin -= dt * ffy * vy[i + nx * (ny + (ny + 1) * k)];
    }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
    {
      bin += dt * ffz * vz[i + nx * j];
      if (overland)
        rain -= dt * ffz * pb->bc_val[5][i + j * nx];
      else
        bin -= dt * ffz * vz[i + nx * (j + ny * nz)];
    }
  if (overland)
  {
    double *qx = (double*)malloc(sizeof(double) * sz);
    double *qy = (double*)malloc(sizeof(double) * sz);
    pfo_overland_kin(pb, p, qx, qy);
    for (int j = 0; j < ny; j++)
      run += dt * dy * (qx[IO2(nx - 1, j)] - qx[IO2(-1, j)]);
    for (int i = 0; i < nx; i++)
      run += dt * dx * (qy[IO2(i, ny - 1)] - qy[IO2(i, -1)]);
    free(qx);
    free(qy);
  }
  if (pb->src)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          src += pb->src[IDX(i, j, k)] * (dx * dy * dz * dt);
  out[0] = bin;
  out[1] = rain;
  out[2] = run;
  out[3] = src;
  free(vx);
  free(vy);
  free(vz);
}

/* ---------------------------------------------------------------------- */
/* Jacobian                                                                */
/* ---------------------------------------------------------------------- */

/* Stencil order of PL/richards_jacobian_eval.c:126-132:
 *   J[0]=C J[1]=W J[2]=E J[3]=S J[4]=N J[5]=L J[6]=U
 * With symm_part (PCMatrixType PFSymmetric) only C,E,N,U are written and
 * the lower couplings are the neighbours' upper ones (PL/matrix.c:258-290,
 * 470-490).  Every array is ghosted (nx+2)(ny+2)(nz+2) and zeroed here.
 *
 * RichardsJacobianEval PL/richards_jacobian_eval.c:217-2200:
 *   time term :445-453; interior faces :594-776; Dirichlet :1049-1226;
 *   FluxBC :1227-1244; OverlandKinematic JC/J :1547-1644, :1938-1987, with
 *   the PFSymmetric+PFMG top-row assembly of PL/pf_hypre.c:360-420 (the top
 *   row keeps the JC centre, surface-surface couplings are dropped). */
void pfo_jacobian_eval(const pfo_problem *pb, const double *p, double dt,
                       int symm_part, double **J)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const long ng = (long)sz * (nz + 2);
  const double dx = pb->dx, dy = pb->dy, dz = pb->dz;
  const double vol = dx * dy * dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double rho = pb->rho, g = pb->gravity, mu = pb->mu;
  const double gc = 0.5 * (g * cos(atan(0.0)) + g * cos(atan(0.0)));
  const double ddp = 0.0;   /* constant density: dRho/dp = 0 */

  for (int s = 0; s < 7; s++)
    if (J[s] && (!symm_part || s == 0 || s == 2 || s == 4 || s == 6))
      memset(J[s], 0, sizeof(double) * ng);

  double *krho = (double*)malloc(sizeof(double) * ng);
  double *dkrho = (double*)malloc(sizeof(double) * ng);
  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        krho[ip] = vg_kr(pb, p[ip], rho, g, pb->alpha[ip], pb->nvg[ip]) * rho;
        dkrho[ip] = vg_dkr(pb, p[ip], rho, g, pb->alpha[ip], pb->nvg[ip]) * rho
                    + vg_kr(pb, p[ip], rho, g, pb->alpha[ip], pb->nvg[ip]) * ddp;
      }

  int overland = (pb->bc[5].type == PFO_BC_OVERLAND);
  double *ke = NULL, *kw = NULL, *kn = NULL, *ks = NULL;
  if (overland)
  {
    ke = (double*)malloc(sizeof(double) * sz);
    kw = (double*)malloc(sizeof(double) * sz);
    kn = (double*)malloc(sizeof(double) * sz);
    ks = (double*)malloc(sizeof(double) * sz);
    pfo_overland_kin_der(pb, p, ke, kw, kn, ks);
  }

  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        double S = vg_sat(p[ip], rho, g, pb->alpha[ip], pb->nvg[ip],
                          pb->sres[ip], pb->ssat[ip]);
        double dS = vg_dsat(p[ip], rho, g, pb->alpha[ip], pb->nvg[ip],
                            pb->sres[ip], pb->ssat[ip]);
        double vol2 = vol * 1.0;
        double c = (dS * rho + S * ddp) * pb->poro[ip] * vol2
                   + pb->ss[ip] * vol2 * (dS * rho * p[ip] + S * ddp * p[ip]
                                         + S * rho);
        double offW = 0, offE = 0, offS = 0, offN = 0, offL = 0, offU = 0;

        /* x faces */
        if (i + 1 < nx)
        {
          int b = ip + 1;
          double diff = p[ip] - p[b];
          double updir = (diff / dx) * gc;
          double x_coeff = 1.0 * dt * ffx * (1.0 / dx) * 1.0
                           * hmean(pb->kx[ip], pb->kx[b]) / mu;
          double sym_west = (-x_coeff * upstream(updir, 0.0, krho[ip], krho[b])) * gc;
          double west_temp = (-x_coeff * diff * upstream(updir, 0.0, dkrho[ip], 0.0)) * gc
                             + sym_west;
          double sym_east = (-x_coeff * upstream(updir, 0.0, krho[ip], krho[b])) * gc;
          double east_temp = (x_coeff * diff * upstream(updir, 0.0, 0.0, dkrho[b])) * gc
                             + sym_east;
          c -= west_temp;
          offE = symm_part ? sym_east : east_temp;
        }
        if (i > 0)
        {
          int a = ip - 1;
          double diff = p[a] - p[ip];
          double updir = (diff / dx) * gc;
          double x_coeff = 1.0 * dt * ffx * (1.0 / dx) * 1.0
                           * hmean(pb->kx[a], pb->kx[ip]) / mu;
          double sym_west = (-x_coeff * upstream(updir, 0.0, krho[a], krho[ip])) * gc;
          double west_temp = (-x_coeff * diff * upstream(updir, 0.0, dkrho[a], 0.0)) * gc
                             + sym_west;
          double sym_east = (-x_coeff * upstream(updir, 0.0, krho[a], krho[ip])) * gc;
          double east_temp = (x_coeff * diff * upstream(updir, 0.0, 0.0, dkrho[ip])) * gc
                             + sym_east;
          c -= east_temp;
          offW = symm_part ? sym_west : west_temp;
        }
        /* y faces */
        if (j + 1 < ny)
        {
          int b = ip + sy;
          double diff = p[ip] - p[b];
          double updir = (diff / dy) * gc;
          double y_coeff = 1.0 * dt * ffy * (1.0 / dy) * 1.0
                           * hmean(pb->ky[ip], pb->ky[b]) / mu;
          double sym_south = -y_coeff * upstream(updir, 0.0, krho[ip], krho[b]) * gc;
          double south_temp = -y_coeff * diff * upstream(updir, 0.0, dkrho[ip], 0.0) * gc
                              + sym_south;
          double sym_north = y_coeff * -upstream(updir, 0.0, krho[ip], krho[b]) * gc;
          double north_temp = y_coeff * diff * upstream(updir, 0.0, 0.0, dkrho[b]) * gc
                              + sym_north;
          c -= south_temp;
          offN = symm_part ? sym_north : north_temp;
        }
        if (j > 0)
        {
          int a = ip - sy;
          double diff = p[a] - p[ip];
          double updir = (diff / dy) * gc;
          double y_coeff = 1.0 * dt * ffy * (1.0 / dy) * 1.0
                           * hmean(pb->ky[a], pb->ky[ip]) / mu;
          double sym_south = -y_coeff * upstream(updir, 0.0, krho[a], krho[ip]) * gc;
          double south_temp = -y_coeff * diff * upstream(updir, 0.0, dkrho[a], 0.0) * gc
                              + sym_south;
          double sym_north = y_coeff * -upstream(updir, 0.0, krho[a], krho[ip]) * gc;
          double north_temp = y_coeff * diff * upstream(updir, 0.0, 0.0, dkrho[ip]) * gc
                              + sym_north;
          c -= north_temp;
          offS = symm_part ? sym_south : south_temp;
        }
        /* z faces */
        if (k + 1 < nz)
        {
          int b = ip + sz;
          double sep = dz * 0.5 * (1.0 + 1.0);
          double lc = p[ip] / sep - (1.0 / (1.0 + 1.0)) * rho * g;
          double uc = p[b] / sep + (1.0 / (1.0 + 1.0)) * rho * g;
          double diff = lc - uc;
          double z_coeff = 1.0 * dt * ffz
                           * hmean_dz(pb->kz[ip], pb->kz[b], 1.0, 1.0) / mu;
          double sym_lower = -z_coeff * (1.0 / (dz * 0.5 * (1.0 + 1.0)))
                             * upstream(lc, uc, krho[ip], krho[b]);
          double lower_temp = -z_coeff * (diff * upstream(lc, uc, dkrho[ip], 0.0)
                                          + (-g * 0.5 * dz * (0.5 * (1.0 + 1.0)) * ddp
                                             * upstream(lc, uc, krho[ip], krho[b])))
                              + sym_lower;
          double sym_upper = z_coeff * (1.0 / (dz * 0.5 * (1.0 + 1.0)))
                             * -upstream(lc, uc, krho[ip], krho[b]);
          c -= lower_temp;
          offU = symm_part ? sym_upper
                           : z_coeff * (diff * upstream(lc, uc, 0.0, dkrho[b])
                                        + (-g * 0.5 * dz * (0.5 * (1.0 + 1.0)) * ddp
                                           * upstream(lc, uc, krho[ip], krho[b])))
                             + sym_upper;
        }
        if (k > 0)
        {
          int a = ip - sz;
          double sep = dz * 0.5 * (1.0 + 1.0);
          double lc = p[a] / sep - (1.0 / (1.0 + 1.0)) * rho * g;
          double uc = p[ip] / sep + (1.0 / (1.0 + 1.0)) * rho * g;
          double diff = lc - uc;
          double z_coeff = 1.0 * dt * ffz
                           * hmean_dz(pb->kz[a], pb->kz[ip], 1.0, 1.0) / mu;
          double sym_lower = -z_coeff * (1.0 / (dz * 0.5 * (1.0 + 1.0)))
                             * upstream(lc, uc, krho[a], krho[ip]);
          double lower_temp = -z_coeff * (diff * upstream(lc, uc, dkrho[a], 0.0)
                                          + (-g * 0.5 * dz * (0.5 * (1.0 + 1.0)) * ddp
                                             * upstream(lc, uc, krho[a], krho[ip])))
                              + sym_lower;
          double sym_upper = z_coeff * (1.0 / (dz * 0.5 * (1.0 + 1.0)))
                             * -upstream(lc, uc, krho[a], krho[ip]);
          double upper_temp = z_coeff * (diff * upstream(lc, uc, 0.0, dkrho[ip])
                                         + (-g * 0.5 * dz * (0.5 * (1.0 + 1.0)) * ddp
                                            * upstream(lc, uc, krho[a], krho[ip])))
                              + sym_upper;
          c -= upper_temp;
          offL = symm_part ? sym_lower : lower_temp;
        }

        /* boundary faces */
        for (int fc = 0; fc < 6; fc++)
        {
          int on = (fc == 0 && i == 0) || (fc == 1 && i == nx - 1) ||
                   (fc == 2 && j == 0) || (fc == 3 && j == ny - 1) ||
                   (fc == 4 && k == 0) || (fc == 5 && k == nz - 1);
          if (!on) continue;
          if (pb->bc[fc].type != PFO_BC_DIR_EQUIL) continue;
          double value = pb->bc_val[fc][fidx(pb, fc, i, j, k)];
          double den_d = rho;
          double prod = krho[ip];
          double prod_der = dkrho[ip];
          double kval = vg_kr(pb, value, rho, g, pb->alpha[ip], pb->nvg[ip]);
          double prod_val = kval * den_d;
          double coeff, diff, o_temp;
          if (fc == 0 || fc == 2)
          {
            double ff = (fc == 0) ? ffx : ffy, dd = (fc == 0) ? dx : dy;
            double kk = (fc == 0) ? pb->kx[ip] : pb->ky[ip];
            coeff = dt * ff * 1.0 * (2.0 / dd) * kk / mu;
            diff = value - p[ip];
            o_temp = coeff * (diff * upstream(value, p[ip], 0.0, prod_der)
                              - upstream(value, p[ip], prod_val, prod));
          }
          else if (fc == 1 || fc == 3)
          {
            double ff = (fc == 1) ? ffx : ffy, dd = (fc == 1) ? dx : dy;
            double kk = (fc == 1) ? pb->kx[ip] : pb->ky[ip];
            coeff = dt * ff * 1.0 * (2.0 / dd) * kk / mu;
            diff = p[ip] - value;
            o_temp = -coeff * (diff * upstream(p[ip], value, prod_der, 0.0)
                               + upstream(p[ip], value, prod, prod_val));
          }
          else if (fc == 4)
          {
            coeff = dt * ffz * (2.0 / (dz * 0.5 * (1.0 + 1.0))) * pb->kz[ip] / mu;
            double lc = (value) - 0.5 * dz * 1.0 * den_d * g;
            double uc = (p[ip]) + 0.5 * dz * 1.0 * rho * g;
            diff = lc - uc;
            o_temp = coeff * (diff * upstream(lc, uc, 0.0, prod_der)
                              + ((-1.0 - g * 0.5 * dz * 1.0 * ddp)
                                 * upstream(lc, uc, prod_val, prod)));
          }
          else
          {
            coeff = dt * ffz * (2.0 / (dz * 0.5 * (1.0 + 1.0))) * pb->kz[ip] / mu;
            double lc = (p[ip]) - 0.5 * dz * 1.0 * rho * g;
            double uc = (value) + 0.5 * dz * 1.0 * den_d * g;
            diff = lc - uc;
            o_temp = -coeff * (diff * upstream(lc, uc, prod_der, 0.0)
                               + ((1.0 - g * 0.5 * dz * 1.0 * ddp)
                                  * upstream(lc, uc, prod, prod_val)));
          }
          c -= o_temp;
        }

        /* kinematic overland top row */
        if (overland && k == nz - 1)
        {
          int io = IO2(i, j);
          double dvol = 0.0;
          if (p[ip] > 0.0)
            dvol = (vol / dz) + (vol / ffy) * dt * (ke[io] - kw[io])
                   + (vol / ffx) * dt * (kn[io] - ks[io]);
          c += dvol;
          if (symm_part)
          {
            /* surface-surface couplings removed from the symmetric PC */
            if (i + 1 < nx) offE = 0.0;
            if (j + 1 < ny) offN = 0.0;
          }
          else
          {
            offW -= (vol / ffy) * dt * (ke[io - 1]);
            offE += (vol / ffy) * dt * (kw[io + 1]);
            offS -= (vol / ffx) * dt * (kn[io - sy]);
            offN += (vol / ffx) * dt * (ks[io + sy]);
            if (i == 0) offW = 0.0;
            if (i == nx - 1) offE = 0.0;
            if (j == 0) offS = 0.0;
            if (j == ny - 1) offN = 0.0;
          }
        }

        J[0][ip] = c;
        J[2][ip] = offE;
        J[4][ip] = offN;
        J[6][ip] = offU;
        if (!symm_part)
        {
          J[1][ip] = offW;
          J[3][ip] = offS;
          J[5][ip] = offL;
        }
      }

  free(krho);
  free(dkrho);
  if (overland) { free(ke); free(kw); free(kn); free(ks); }
}

/* y = J x on the 7-point stencil (PL/matvec.c:40-330; symmetric storage
 * reads the lower couplings from the neighbour's upper coefficient). */
void pfo_matvec(const pfo_problem *pb, double **J, int symm, const double *x,
                double *y)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  #pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        double w = symm ? J[2][ip - 1] : J[1][ip];
        double s = symm ? J[4][ip - sy] : J[3][ip];
        double l = symm ? J[6][ip - sz] : J[5][ip];
        double v = J[0][ip] * x[ip];
        if (i > 0) v += w * x[ip - 1];
        if (i + 1 < nx) v += J[2][ip] * x[ip + 1];
        if (j > 0) v += s * x[ip - sy];
        if (j + 1 < ny) v += J[4][ip] * x[ip + sy];
        if (k > 0) v += l * x[ip - sz];
        if (k + 1 < nz) v += J[6][ip] * x[ip + sz];
        y[ip] = v;
      }
}

/* ---------------------------------------------------------------------- */
/* vector utilities (PL/vector_utilities.c, interior cells only)          */
/* ---------------------------------------------------------------------- */

double pfo_dot(const pfo_problem *pb, const double *x, const double *y)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  double s = 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        s += x[ip] * y[ip];
      }
  return s;
}

double pfo_l2norm(const pfo_problem *pb, const double *x)
{
  return sqrt(pfo_dot(pb, x, x));
}

double pfo_maxnorm(const pfo_problem *pb, const double *x)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  double m = 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        double v = fabs(x[IDX(i, j, k)]);
        if (v > m) m = v;
      }
  return m;
}

double pfo_l1norm(const pfo_problem *pb, const double *x)
{
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  double s = 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        s += fabs(x[IDX(i, j, k)]);
  return s;
}

/* Subsurface storage sum(phi*S*V + [p >= 0] p*S*Ss*V) and surface storage
 * sum(max(p_top*dx*dy, 0)): pftools/python/parflow/tools/hydrology.py
 * :136-205 restated for a box. */
void pfo_storage(const pfo_problem *pb, const double *p, const double *s,
                 double *subsurface, double *surface)
{
  /* calculate_subsurface_storage / calculate_surface_storage,
   * pftools/python/parflow/tools/hydrology.py:136-205: per cell
   * incompressible = porosity*saturation*dz*dx*dy, compressible =
   * pressure*saturation*specific_storage*dz*dx*dy set to 0 where pressure < 0;
   * per top cell pressure*dx*dy set to 0 where negative; summed x fastest. */
  const int nx = pb->nx, ny = pb->ny, nz = pb->nz;
  const int sy = nx + 2, sz = (nx + 2) * (ny + 2);
  const double dx = pb->dx, dy = pb->dy, dz = pb->dz;
  double sub = 0.0, sur = 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        int ip = IDX(i, j, k);
        double inc = pb->poro[ip] * s[ip] * dz * dx * dy;
        double comp = p[ip] * s[ip] * pb->ss[ip] * dz * dx * dy;
        if (p[ip] < 0.0) comp = 0.0;
        sub += inc + comp;
      }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
    {
      double t = p[IDX(i, j, nz - 1)] * dx * dy;
      if (t < 0.0) t = 0.0;
      sur += t;
    }
  *subsurface = sub;
  *surface = sur;
}
