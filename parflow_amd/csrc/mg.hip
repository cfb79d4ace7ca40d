// mg.hip -- semicoarsening multigrid preconditioner for gfx950.
//
// Structure of ParFlow's in-repo MGSemi (PL/mg_semi.c): each level halves
// the direction of smallest spacing that still has >= 2 points
// (:832-900), coarse points are the even indices, prolongation weights are
// operator induced at the odd (fine) points and coarse operators are the
// collapsed Galerkin product (SetupCoarseOps :437-760); transfers follow
// PL/mg_semi_restrict.c / PL/mg_semi_prolong.c and the smoother is point
// red/black Gauss-Seidel (PL/rb_GS_point.c, red = (i+j+k) odd).
//
// One template serves FP64 and FP32 level storage (the mixed-precision
// preconditioner keeps every level, vector and operator in FP32; only the
// interface vectors are FP64).  Matrices are symmetric-stored C,E,N,U.
//
// MI355X mapping:
//  * every per-point operation is a device function; large levels run them
//    as one-thread-per-point x-coalesced grid kernels whose tiles are
//    renumbered so each of the 8 XCDs owns a contiguous z-range (the
//    hardware deals workgroups to XCDs round-robin; without the renumbering
//    the z-neighbour planes of a tile sit in other XCDs' L2);
//  * every level with <= MG_SMALL_DEFAULT points (the bottom of the V-cycle,
//    a dozen levels of a semicoarsened hierarchy; PFHIP_MG_SMALL overrides
//    the bound) runs inside ONE workgroup of 1024 threads, phase by phase
//    with __syncthreads, instead of ~100 launches that are pure launch
//    latency.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include "pfhip_internal.h"
#include "device_util.h"

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      pfhip_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x,             \
                      hipGetErrorString(e_));                              \
      return PFHIP_ERR_HIP;                                                \
    }                                                                      \
  } while (0)
#define LAUNCHCHK() HIPCHK(hipGetLastError())
#define STREAM(s) ((hipStream_t)(s))
#define MAXL 128
#define MG_SMALL_DEFAULT 4096
#define MG_LDS_DEFAULT (144L * 1024)   // LDS of the bottom-of-cycle workgroup
#define CWG 1024
#define MAX_SWEEPS 64

struct LGeom {
  int n[3];
  int s[3];   // strides: 1, n0+2, (n0+2)(n1+2)
};

static LGeom lgeom(const int *n)
{
  LGeom g;
  for (int q = 0; q < 3; q++) g.n[q] = n[q];
  g.s[0] = 1; g.s[1] = n[0] + 2; g.s[2] = (n[0] + 2) * (n[1] + 2);
  return g;
}
static long lsize(const int *n) { return (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2); }
static long npts(const int *n) { return (long)n[0] * n[1] * n[2]; }

#define LIDX(g, i, j, k) ((i) + 1 + ((j) + 1) * (g).s[1] + ((k) + 1) * (g).s[2])

// ---- XCD-aware tile numbering ---------------------------------------------
// Tiles are 64 x 4 points of one plane; nb tiles in x-fastest order.  The
// launch has nb rounded up to a multiple of 8 workgroups; hardware block b
// runs on XCD b % 8, so logical tile (b % 8) * (nb8 / 8) + b / 8 gives XCD x
// the contiguous tile range [x * nb8/8, (x+1) * nb8/8).
static inline unsigned ntiles(const LGeom &g)
{
  unsigned nb = (unsigned)((g.n[0] + 63) / 64) * ((g.n[1] + 3) / 4) * g.n[2];
  return (nb + 7u) & ~7u;
}
#define LBLOCK dim3(64, 4, 1)

__device__ __forceinline__ bool xcd_point(const LGeom &g, int &i, int &j, int &k)
{
  int gx = (g.n[0] + 63) >> 6, gy = (g.n[1] + 3) >> 2;
  int nb = gx * gy * g.n[2];
  int nb8 = (nb + 7) & ~7;
  int b = blockIdx.x;
  int L = (b & 7) * (nb8 >> 3) + (b >> 3);
  if (L >= nb) return false;
  int bx = L % gx, r = L / gx;
  int by = r % gy, bz = r / gy;
  i = bx * 64 + threadIdx.x;
  j = by * 4 + threadIdx.y;
  k = bz;
  return i < g.n[0] && j < g.n[1];
}

#define XCD_IJK(g)                    \
  int i, j, k;                        \
  if (!xcd_point((g), i, j, k)) return;

__device__ __forceinline__ void unflat(const LGeom &g, int idx, int &i, int &j, int &k)
{
  i = idx % g.n[0];
  int r = idx / g.n[0];
  j = r % g.n[1];
  k = r / g.n[1];
}

// C, upper E, N, U and, for full (non-symmetric) storage, lower W, S, L;
// with symmetric storage l[] is null and a point's lower coupling is its
// lower neighbour's upper one
template <typename T> struct Mat { T *c, *u[3], *l[3]; };

template <typename T> struct LevelDev {
  LGeom g;
  int dir, o1, o2;
  Mat<T> A;
  T *x, *b, *t, *p1, *p2;
};

// coupling of point ip to its lower neighbour in direction q (stride s)
template <typename T, bool FULL>
__device__ __forceinline__ T lo_cpl(const Mat<T> &A, long ip, int q, long s)
{
  return FULL ? A.l[q][ip] : A.u[q][ip - s];
}

// coupling of point ip in direction q on side -1/+1
template <typename T, bool FULL>
__device__ __forceinline__ T cpl(const Mat<T> &A, const LGeom &g, long ip, int q, int side)
{
  return (side > 0) ? A.u[q][ip] : lo_cpl<T, FULL>(A, ip, q, g.s[q]);
}

// ---- per-point operations -------------------------------------------------
// Coarse operator of one coarse point (SetupCoarseOps, PL/mg_semi.c:437-760)
// in one pass: the operator-induced prolongation weights of the odd fine
// points beside it (:505-545; the one above is stored for the transfers,
// the one below is recomputed), the collapsed Galerkin couplings (:560-700)
// and the centre completion (:720-750), whose lower-neighbour couplings are
// recomputed from the fine operator instead of read back, with the
// reference's operand order.
//
// With full storage (the FullJacobian PC matrix) each side of a row is
// collapsed from its own couplings: along d the coarse couplings are the
// P^T A P ones, a_U(C) p2(C+1) and a_L(C) p1(C-1), and the lower off-
// direction couplings are collapsed like the upper ones; on a symmetric
// matrix this gives the symmetric formulas.
template <typename T, bool FULL>
__device__ __forceinline__ void pw_calc(const LGeom &g, const Mat<T> &A, int d, int o1, int o2,
                                        long ip, T &w1, T &w2)
{
  T a0 = A.c[ip];
  T a1 = cpl<T, FULL>(A, g, ip, d, -1), a2 = cpl<T, FULL>(A, g, ip, d, +1);
  T a3 = cpl<T, FULL>(A, g, ip, o1, -1), a4 = cpl<T, FULL>(A, g, ip, o1, +1);
  T a5 = cpl<T, FULL>(A, g, ip, o2, -1), a6 = cpl<T, FULL>(A, g, ip, o2, +1);
  T ap0 = a0 + a3 + a4 + a5 + a6;
  if (ap0 != T(0)) { w1 = -a1 / ap0; w2 = -a2 / ap0; }
  else { w1 = T(0); w2 = T(0); }
}

template <typename T>
__device__ __forceinline__ T collapse3(const T *a, long iA, long sd)
{
  return a[iA] + T(0.5) * a[iA - sd] + T(0.5) * a[iA + sd];
}

template <typename T>
__device__ __forceinline__ T up_coupling(const Mat<T> &A, int q, long iA, long sd)
{
  return collapse3(A.u[q], iA, sd);
}

template <typename T, bool FULL>
__device__ __forceinline__ void coarse_point(const LGeom &gf, const LGeom &gc, const Mat<T> &A,
                                             const Mat<T> &Ac, int d, int o1, int o2, T *p1,
                                             T *p2, int i, int j, int k)
{
  int cid[3] = { i, j, k };
  int fi[3] = { i, j, k };
  fi[d] *= 2;
  long iA = LIDX(gf, fi[0], fi[1], fi[2]);
  long iAc = LIDX(gc, i, j, k);
  long sd = gf.s[d];
  long iA1 = iA - sd, iA2 = iA + sd;
  // weights of the odd point below (owned by the coarse point below) and
  // above (owned here); absent points have zero weights
  T p2lo = T(0), p1lo = T(0), p1hi = T(0), p2hi = T(0);
  if (fi[d] > 0) {
    T w1, w2;
    pw_calc<T, FULL>(gf, A, d, o1, o2, iA1, w1, w2);
    p2lo = w2; p1lo = w1;
  }
  if (fi[d] + 1 < gf.n[d]) {
    T w1, w2;
    pw_calc<T, FULL>(gf, A, d, o1, o2, iA2, w1, w2);
    p1[iA2] = w1; p2[iA2] = w2;
    p1hi = w1; p2hi = w2;
  }
  T up1 = up_coupling(A, o1, iA, sd);
  T up2 = up_coupling(A, o2, iA, sd);
  T c4 = (cid[o1] == gc.n[o1] - 1) ? T(0) : up1;
  T c6 = (cid[o2] == gc.n[o2] - 1) ? T(0) : up2;
  Ac.u[o1][iAc] = c4;
  Ac.u[o2][iAc] = c6;
  T c3, c5;
  if (FULL) {
    c3 = (cid[o1] > 0) ? collapse3(A.l[o1], iA, sd) : T(0);
    c5 = (cid[o2] > 0) ? collapse3(A.l[o2], iA, sd) : T(0);
    Ac.l[o1][iAc] = c3;
    Ac.l[o2][iAc] = c5;
    Ac.u[d][iAc] = (cid[d] + 1 < gc.n[d]) ? A.u[d][iA] * p2hi : T(0);
    Ac.l[d][iAc] = (cid[d] > 0) ? A.l[d][iA] * p1lo : T(0);
  } else {
    T upd;
    if (cid[d] + 1 < gc.n[d]) upd = cpl<T, false>(A, gf, iA + 2 * sd, d, -1) * p1hi;
    else upd = T(0);
    Ac.u[d][iAc] = upd;
    // lower neighbours' o1 / o2 couplings (never the last index, so unzeroed)
    c3 = (cid[o1] > 0) ? up_coupling(A, o1, iA - gf.s[o1], sd) : T(0);
    c5 = (cid[o2] > 0) ? up_coupling(A, o2, iA - gf.s[o2], sd) : T(0);
  }
  T a0 = A.c[iA];
  T a1 = cpl<T, FULL>(A, gf, iA, d, -1), a2 = cpl<T, FULL>(A, gf, iA, d, +1);
  T a3 = cpl<T, FULL>(A, gf, iA, o1, -1), a4 = cpl<T, FULL>(A, gf, iA, o1, +1);
  T a5 = cpl<T, FULL>(A, gf, iA, o2, -1), a6 = cpl<T, FULL>(A, gf, iA, o2, +1);
  T cc = a0 + a3 + a4 + a5 + a6 + a1 * p2lo + a2 * p1hi;
  Ac.c[iAc] = cc - (c3 + c4 + c5 + c6);
}

// The smoother's division by the diagonal: IEEE division in FP64; in FP32
// the hardware reciprocal with one Newton correction of the quotient
// (4 instructions instead of the 11 of IEEE division, within 1 ulp for the
// normal, finite operands a diagonal and a residual are).  Every FP32
// smoothing path divides through this, so they all give the same bits.
template <typename T> __device__ __forceinline__ T gs_div(T num, T c) { return num / c; }
template <> __device__ __forceinline__ float gs_div<float>(float num, float c)
{
  const float r = __builtin_amdgcn_rcpf(c);
  const float q = num * r;
  return __builtin_fmaf(__builtin_fmaf(-c, q, num), r, q);
}

// Gauss-Seidel point update x = (b - sum of present neighbour terms) / c,
// terms summed west, east, south, north, lower, upper (PL/rb_GS_point.c:
// 200-290); absent neighbours contribute exactly zero
template <typename T>
__device__ __forceinline__ T gs_update(T bb, T c, T aw, T xw, bool hw, T ae, T xe, bool he,
                                       T as, T xs, bool hs, T an, T xn, bool hn, T al, T xl,
                                       bool hl, T au, T xu, bool hu)
{
  T w = hw ? aw * xw : T(0);
  T e = he ? ae * xe : T(0);
  T so = hs ? as * xs : T(0);
  T no = hn ? an * xn : T(0);
  T lo = hl ? al * xl : T(0);
  T up = hu ? au * xu : T(0);
  T sum = w + e + so + no + lo + up;
  return gs_div(bb - sum, c);
}

// red/black update of one point of parity `par`; init: x = b / a0
template <typename T, bool FULL>
__device__ __forceinline__ void rbgs_point(const LGeom &g, const Mat<T> &A, T *x, const T *b,
                                           int par, int init, int i, int j, int k)
{
  if (((i + j + k) & 1) != par) return;
  long ip = LIDX(g, i, j, k);
  if (init) { x[ip] = gs_div(b[ip], A.c[ip]); return; }
  const int sy = g.s[1], sz = g.s[2];
  bool hw = i > 0, he = i + 1 < g.n[0], hs = j > 0, hn = j + 1 < g.n[1];
  bool hl = k > 0, hu = k + 1 < g.n[2];
  x[ip] = gs_update(b[ip], A.c[ip],
                    hw ? lo_cpl<T, FULL>(A, ip, 0, 1) : T(0), hw ? x[ip - 1] : T(0), hw,
                    A.u[0][ip], he ? x[ip + 1] : T(0), he,
                    hs ? lo_cpl<T, FULL>(A, ip, 1, sy) : T(0), hs ? x[ip - sy] : T(0), hs,
                    A.u[1][ip], hn ? x[ip + sy] : T(0), hn,
                    hl ? lo_cpl<T, FULL>(A, ip, 2, sz) : T(0), hl ? x[ip - sz] : T(0), hl,
                    A.u[2][ip], hu ? x[ip + sz] : T(0), hu);
}

template <typename T, bool FULL>
__device__ __forceinline__ void resid_point(const LGeom &g, const Mat<T> &A, const T *x,
                                            const T *b, T *t, int i, int j, int k)
{
  long ip = LIDX(g, i, j, k);
  const int sy = g.s[1], sz = g.s[2];
  T v = A.c[ip] * x[ip];
  if (i > 0) v += lo_cpl<T, FULL>(A, ip, 0, 1) * x[ip - 1];
  if (i + 1 < g.n[0]) v += A.u[0][ip] * x[ip + 1];
  if (j > 0) v += lo_cpl<T, FULL>(A, ip, 1, sy) * x[ip - sy];
  if (j + 1 < g.n[1]) v += A.u[1][ip] * x[ip + sy];
  if (k > 0) v += lo_cpl<T, FULL>(A, ip, 2, sz) * x[ip - sz];
  if (k + 1 < g.n[2]) v += A.u[2][ip] * x[ip + sz];
  t[ip] = b[ip] - v;
}

// resid_point's value without storing it
template <typename T, bool FULL>
__device__ __forceinline__ T resid_value(const LGeom &g, const Mat<T> &A, const T *x, const T *b,
                                         int i, int j, int k)
{
  long ip = LIDX(g, i, j, k);
  const int sy = g.s[1], sz = g.s[2];
  T v = A.c[ip] * x[ip];
  if (i > 0) v += lo_cpl<T, FULL>(A, ip, 0, 1) * x[ip - 1];
  if (i + 1 < g.n[0]) v += A.u[0][ip] * x[ip + 1];
  if (j > 0) v += lo_cpl<T, FULL>(A, ip, 1, sy) * x[ip - sy];
  if (j + 1 < g.n[1]) v += A.u[1][ip] * x[ip + sy];
  if (k > 0) v += lo_cpl<T, FULL>(A, ip, 2, sz) * x[ip - sz];
  if (k + 1 < g.n[2]) v += A.u[2][ip] * x[ip + sz];
  return b[ip] - v;
}

// r_c = r_f[C] + p2[C-1] r_f[C-1] + p1[C+1] r_f[C+1]  (mg_semi_restrict.c:120-150)
template <typename T>
__device__ __forceinline__ void restrict_point(const LGeom &gf, const LGeom &gc, int d,
                                               const T *rf, T *rc, const T *p1, const T *p2,
                                               int i, int j, int k)
{
  int fi[3] = { i, j, k };
  fi[d] *= 2;
  long iF = LIDX(gf, fi[0], fi[1], fi[2]);
  long sd = gf.s[d];
  T lo = (fi[d] > 0) ? p2[iF - sd] * rf[iF - sd] : T(0);
  T hi = (fi[d] + 1 < gf.n[d]) ? p1[iF + sd] * rf[iF + sd] : T(0);
  rc[LIDX(gc, i, j, k)] = rf[iF] + (lo + hi);
}

// x_f += P e_c (injection at coarse points, interpolation at fine points;
// mg_semi_prolong.c:85-160)
template <typename T>
__device__ __forceinline__ void prolong_point(const LGeom &gf, const LGeom &gc, int d,
                                              const T *ec, T *xf, const T *p1, const T *p2,
                                              int i, int j, int k)
{
  int id[3] = { i, j, k };
  long ip = LIDX(gf, i, j, k);
  int c[3] = { i, j, k };
  c[d] = id[d] >> 1;
  if ((id[d] & 1) == 0) {
    xf[ip] += ec[LIDX(gc, c[0], c[1], c[2])];
    return;
  }
  T lo = p1[ip] * ec[LIDX(gc, c[0], c[1], c[2])];
  T hi = T(0);
  if (id[d] + 1 < gf.n[d]) {
    c[d] = (id[d] + 1) >> 1;
    hi = p2[ip] * ec[LIDX(gc, c[0], c[1], c[2])];
  }
  xf[ip] += lo + hi;
}

// ---- grid kernels (large levels) --------------------------------------------
template <typename T, bool FULL>
__global__ __launch_bounds__(256) void k_coarse_fused(LGeom gf, LGeom gc, Mat<T> A, Mat<T> Ac,
                                                      int d, int o1, int o2, T *__restrict__ p1,
                                                      T *__restrict__ p2)
{
  XCD_IJK(gc);
  coarse_point<T, FULL>(gf, gc, A, Ac, d, o1, o2, p1, p2, i, j, k);
}

// Coarse operators of a z-semicoarsened level (symmetric storage) in one
// z-marching pass: a thread owns an (i, j) column of 64 x 4 column tiles
// and walks the coarse planes upwards.  The odd fine plane between coarse
// points C and C+2 is read once, as the plane above C (its weights and
// couplings), and its values are carried as the plane below C+2, so a
// coarse point loads two fine planes (12 values) instead of coarse_point's
// four.  The next point's loads are issued before the current point is
// computed.  The expressions and their order are coarse_point's for d = 2
// (o1 = y, o2 = x), so the result is the same.
template <typename T> struct CzPlane { T c, ux, uxw, uy, uys, uz; };

template <typename T>
__device__ __forceinline__ void cz_load(CzPlane<T> &P, const Mat<T> &A, long ip, long sy)
{
  P.c = A.c[ip];
  P.ux = A.u[0][ip]; P.uxw = A.u[0][ip - 1];
  P.uy = A.u[1][ip]; P.uys = A.u[1][ip - sy];
  P.uz = A.u[2][ip];
}

template <typename T>
__global__ __launch_bounds__(256) void k_coarse_z(LGeom gf, LGeom gc, Mat<T> A, Mat<T> Ac,
                                                  T *__restrict__ p1, T *__restrict__ p2)
{
  const int gx = (gc.n[0] + 63) >> 6, gy = (gc.n[1] + 3) >> 2;
  const int nb = gx * gy, nb8 = (nb + 7) & ~7;
  const int b = blockIdx.x;
  const int L = (b & 7) * (nb8 >> 3) + (b >> 3);
  if (L >= nb) return;
  const int i = (L % gx) * 64 + threadIdx.x, j = (L / gx) * 4 + threadIdx.y;
  if (i >= gc.n[0] || j >= gc.n[1]) return;
  const long sy = gf.s[1], sz = gf.s[2];
  const int nfz = gf.n[2], ncz = gc.n[2];
  // values of the odd plane below (the ghost plane below the first point:
  // every level array's ghost entries are zero)
  T lo_uy = T(0), lo_uys = T(0), lo_ux = T(0), lo_uxw = T(0), lo_uz = T(0), p2lo = T(0);
  long iA = LIDX(gf, i, j, 0);
  long iAc = LIDX(gc, i, j, 0);
  CzPlane<T> e, o, en, on;
  cz_load(e, A, iA, sy);
  cz_load(o, A, iA + sz, sy);            // the ghost plane above when nfz == 1
  for (int kc = 0; kc < ncz; kc++) {
    const int fk = 2 * kc;
    if (kc + 1 < ncz) {
      cz_load(en, A, iA + 2 * sz, sy);
      cz_load(on, A, iA + 3 * sz, sy);   // at most the ghost plane above
    }
    const bool hi = fk + 1 < nfz;
    // weights of the odd point above (pw_calc)
    T w1 = T(0), w2 = T(0);
    if (hi) {
      T ap0 = o.c + o.uys + o.uy + o.uxw + o.ux;
      if (ap0 != T(0)) { w1 = -e.uz / ap0; w2 = -o.uz / ap0; }
      p1[iA + sz] = w1; p2[iA + sz] = w2;
    }
    const T p1hi = w1;
    T up1 = e.uy + T(0.5) * lo_uy + T(0.5) * o.uy;
    T up2 = e.ux + T(0.5) * lo_ux + T(0.5) * o.ux;
    T c4 = (j == gc.n[1] - 1) ? T(0) : up1;
    T c6 = (i == gc.n[0] - 1) ? T(0) : up2;
    Ac.u[1][iAc] = c4;
    Ac.u[0][iAc] = c6;
    Ac.u[2][iAc] = (kc + 1 < ncz) ? o.uz * p1hi : T(0);
    T cc = e.c + e.uys + e.uy + e.uxw + e.ux + lo_uz * p2lo + e.uz * p1hi;
    T c3 = (j > 0) ? e.uys + T(0.5) * lo_uys + T(0.5) * o.uys : T(0);
    T c5 = (i > 0) ? e.uxw + T(0.5) * lo_uxw + T(0.5) * o.uxw : T(0);
    Ac.c[iAc] = cc - (c3 + c4 + c5 + c6);
    lo_uy = o.uy; lo_uys = o.uys; lo_ux = o.ux; lo_uxw = o.uxw; lo_uz = o.uz;
    p2lo = w2;
    e = en; o = on;
    iA += 2 * sz;
    iAc += gc.s[2];
  }
}

template <typename T, bool FULL>
__global__ __launch_bounds__(256) void k_rbgs(LGeom g, Mat<T> A, T *__restrict__ x,
                                              const T *__restrict__ b, int par, int init)
{
  XCD_IJK(g);
  rbgs_point<T, FULL>(g, A, x, b, par, init, i, j, k);
}

// ---- fused red -> black sweep pair ------------------------------------------
// One launch performs a red half-sweep followed by a black half-sweep with
// the arithmetic of two k_rbgs launches.  A workgroup owns PX x TY columns
// and marches a chunk of KC planes; its threads cover one extra ring column
// on each side, whose red values it recomputes so that the black update at
// the tile edge sees the new red neighbours.  Plane k: red at k (into an LDS
// ring of 3 planes), barrier, black at k-1 from the red planes k-2, k-1, k.
// Reads x_in (old values) and writes x_out (every interior point exactly
// once), so workgroups never read what another one writes.  Traffic per
// point: C, E, N, U, b and x_in read, x_out written (7 words instead of 13).
#define PX 62

// everything one point's update reads: centre, b, the six couplings and the
// six neighbour values of x_in.  Every load is issued unconditionally from a
// clamped in-bounds point (ghost entries of every level array are 0 and
// absent neighbours are masked by gs_update), so the loop body is straight-
// line code and the prefetch of the next plane stays in flight while the
// current one is computed.
template <typename T> struct PointSet { T c, bb, aw, ae, as, an, al, au, xw, xe, xs, xn, xl, xu; };

template <typename T, bool INIT>
__device__ __forceinline__ void load_point(PointSet<T> &S, const Mat<T> &A,
                                           const T *__restrict__ xin,
                                           const T *__restrict__ b, unsigned ip, unsigned sy,
                                           unsigned sz)
{
  S.c = A.c[ip]; S.bb = b[ip];
  S.aw = A.u[0][ip - 1]; S.ae = A.u[0][ip];
  S.as = A.u[1][ip - sy]; S.an = A.u[1][ip];
  S.al = A.u[2][ip - sz]; S.au = A.u[2][ip];
  if (INIT) {
    S.xw = S.xe = S.xs = S.xn = S.xl = S.xu = T(0);
  } else {
    S.xw = xin[ip - 1]; S.xe = xin[ip + 1];
    S.xs = xin[ip - sy]; S.xn = xin[ip + sy];
    S.xl = xin[ip - sz]; S.xu = xin[ip + sz];
  }
}

// state a column carries from plane to plane
template <typename T> struct ColState {
  T sc, sb, sw, se, ss, sn, sl, su;   // coefficients of the last black point
  T rm1, rm2;                         // red values at k-1, k-2
};

// one plane of the march: red at k from `cur` (loaded one step earlier),
// black at k-1, one store per point of plane k-1 (its red value computed
// last step or its black value computed now).  Branch-free apart from the
// store.
template <typename T, int TY, bool INIT>
__device__ __forceinline__ void pair_step(T (*red)[TY + 2][64], int slot, int pslot,
                                          const PointSet<T> &cur, ColState<T> &cs, int k,
                                          int kb, int ke, int nz, bool valid, bool owned,
                                          bool hw, bool he, bool hs, bool hn, int i, int j,
                                          T *__restrict__ xout, unsigned ip, unsigned sz)
{
  const int tx = threadIdx.x, ty = threadIdx.y;
  const bool isred = ((i + j + k) & 1) == 1;
  const bool here = valid && k >= 0 && k < nz;
  T rr = INIT ? gs_div(cur.bb, cur.c)
              : gs_update(cur.bb, cur.c, cur.aw, cur.xw, hw, cur.ae, cur.xe, he, cur.as,
                          cur.xs, hs, cur.an, cur.xn, hn, cur.al, cur.xl, k > 0, cur.au,
                          cur.xu, k + 1 < nz);
  const T rk = (here && isred) ? rr : T(0);
  const bool keep = here && !isred;
  cs.sc = keep ? cur.c : cs.sc;   cs.sb = keep ? cur.bb : cs.sb;
  cs.sw = keep ? cur.aw : cs.sw;  cs.se = keep ? cur.ae : cs.se;
  cs.ss = keep ? cur.as : cs.ss;  cs.sn = keep ? cur.an : cs.sn;
  cs.sl = keep ? cur.al : cs.sl;  cs.su = keep ? cur.au : cs.su;
  red[slot][ty][tx] = rk;
  // black at k-1 (black exactly when (i,j,k) is red); in-plane red
  // neighbours from the plane stored last step (ring lanes read clamped
  // slots and never store)
  const T(*R)[64] = red[pslot];
  const int txm = tx > 0 ? tx - 1 : 0, txp = tx < 63 ? tx + 1 : 63;
  const int tym = ty > 0 ? ty - 1 : 0, typ = ty < TY + 1 ? ty + 1 : TY + 1;
  const int kk = k - 1;
  T vb = gs_update(cs.sb, cs.sc, cs.sw, R[ty][txm], hw, cs.se, R[ty][txp], he,
                   cs.ss, R[tym][tx], hs, cs.sn, R[typ][tx], hn,
                   cs.sl, cs.rm2, kk > 0, cs.su, rk, kk + 1 < nz);
  T v = isred ? vb : cs.rm1;
  if (owned && kk >= kb && kk < ke) xout[ip - sz] = v;
  cs.rm2 = cs.rm1;
  cs.rm1 = rk;
}

template <typename T, int TY, int KC, bool INIT>
__global__ __launch_bounds__(64 * (TY + 2)) void k_rbgs_pair(LGeom g, Mat<T> A,
                                                             const T *__restrict__ xin,
                                                             T *__restrict__ xout,
                                                             const T *__restrict__ b)
{
  __shared__ T red[3][TY + 2][64];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = (g.n[0] + PX - 1) / PX, gy = (g.n[1] + TY - 1) / TY;
  const int gz = (g.n[2] + KC - 1) / KC;
  const int nb = gx * gy * gz, nb8 = (nb + 7) & ~7;
  const int bid = blockIdx.x;
  const int Lt = (bid & 7) * (nb8 >> 3) + (bid >> 3);
  if (Lt >= nb) return;                       // uniform over the workgroup
  const int bx = Lt % gx, rr = Lt / gx;
  const int by = rr % gy, bz = rr / gy;
  const int i = bx * PX - 1 + tx, j = by * TY - 1 + ty;
  const bool valid = i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1];
  const bool owned = valid && tx >= 1 && tx <= PX && ty >= 1 && ty <= TY;
  const int nz = g.n[2];
  const int kb = bz * KC, ke = min(kb + KC, nz);
  // 32-bit unsigned offsets: every level array has < 2^32 entries
  const unsigned sy = g.s[1], sz = g.s[2];
  const bool hw = i > 0, he = i + 1 < g.n[0], hs = j > 0, hn = j + 1 < g.n[1];
  // loads come from the clamped column (ic, jc) and planes in [0, nz-1]
  const int ic = min(max(i, 0), g.n[0] - 1), jc = min(max(j, 0), g.n[1] - 1);
  const unsigned col = (unsigned)LIDX(g, ic, jc, 0) - sz;   // plane k at col + (k+1) sz
  unsigned ip = valid ? (unsigned)LIDX(g, i, j, kb - 1) : 0u; // store address, plane kb-1
  ColState<T> cs;
  cs.sc = T(1); cs.sb = cs.sw = cs.se = cs.ss = cs.sn = cs.sl = cs.su = T(0);
  cs.rm1 = cs.rm2 = T(0);
  PointSet<T> P0, P1;                         // two register sets, roles alternate
  load_point<T, INIT>(P0, A, xin, b, col + (unsigned)(max(kb - 1, 0) + 1) * sz, sy, sz);
  int it = 0;
  for (int k = kb - 1; k <= ke; k += 2) {
    // plane k from P0, prefetch k+1 into P1
    load_point<T, INIT>(P1, A, xin, b, col + (unsigned)(min(k + 1, nz - 1) + 1) * sz, sy, sz);
    pair_step<T, TY, INIT>(red, it % 3, (it + 2) % 3, P0, cs, k, kb, ke, nz, valid, owned, hw,
                           he, hs, hn, i, j, xout, ip, sz);
    lds_barrier();
    it++; ip += sz;
    if (k + 1 > ke) break;
    // plane k+1 from P1, prefetch k+2 into P0
    load_point<T, INIT>(P0, A, xin, b, col + (unsigned)(min(k + 2, nz - 1) + 1) * sz, sy, sz);
    pair_step<T, TY, INIT>(red, it % 3, (it + 2) % 3, P1, cs, k + 1, kb, ke, nz, valid, owned,
                           hw, he, hs, hn, i, j, xout, ip, sz);
    lds_barrier();
    it++; ip += sz;
  }
}

// ---- fused pair, two points per lane ----------------------------------------
// The same red -> black pair for levels with an even x extent, where every
// lane owns the x-adjacent points (i0, i0+1), i0 even: each load fetches both
// points as one 8/16-byte vector (the level arrays start one element past
// an aligned address, so the ghosted index of an even i0 is aligned), the
// x-neighbours come from the lanes beside through lane exchanges, the z
// neighbours and the lower coupling are carried from plane to plane, so a
// plane costs 9 vector loads per lane for 2 points instead of 14 per point
// (the vector memory pipeline was the limiter).  A row of a tile is one
// wavefront: lanes 1..62 own 124 points, lane 0's second and lane 63's
// first point are the ring whose red values the tile recomputes.
#define PX2 124

template <typename T> struct V2T;
template <> struct V2T<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct V2T<double> { typedef double type __attribute__((ext_vector_type(2))); };

// lane l receives lane l-1's value (lane 0 its own), lane l+1's (lane 63
// its own): bpermute lane exchanges within the wavefront
template <typename T> __device__ __forceinline__ T from_prev(T v) { return __shfl_up(v, 1, 64); }
template <typename T> __device__ __forceinline__ T from_next(T v) { return __shfl_down(v, 1, 64); }
// sum over the wavefront's 64 lanes (butterfly; lane 0's value is used)
__device__ __forceinline__ double wave_sum_d(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// element ip of an array through a 32-bit byte offset from its base (a
// level array is below 4 GiB: problem creation bounds the finest grid), so
// the access takes the base from scalar registers and one 32-bit vector
// offset instead of a 64-bit address per lane
template <typename V, typename T>
__device__ __forceinline__ V ldo(const T *__restrict__ base, unsigned ip)
{
  return *(const V *)((const char *)base + ip * (unsigned)sizeof(T));
}
template <typename V, typename T>
__device__ __forceinline__ void sto(T *__restrict__ base, unsigned ip, V v)
{
  *(V *)((char *)base + ip * (unsigned)sizeof(T)) = v;
}
// the same for device-memory arrays whose address the kernel reads from
// memory (interface vectors of a replayed launch sequence): global-segment
// accesses instead of flat ones
template <typename T>
__device__ __forceinline__ T ldg_o(const T *base, unsigned ip)
{
  return *(const __attribute__((address_space(1))) T *)((const char *)base +
                                                         ip * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void stg_o(T *base, unsigned ip, T v)
{
  *(__attribute__((address_space(1))) T *)((char *)base + ip * (unsigned)sizeof(T)) = v;
}

// 16-byte FP64 pair access for interface vectors whose even-i points sit on
// a 16-byte boundary (the solver's vectors that meet the multigrid start one
// element past an aligned address, as the level arrays do)
typedef V2T<double>::type D2;
__device__ __forceinline__ D2 ldg_o2(const double *base, unsigned ip)
{
  return *(const __attribute__((address_space(1))) D2 *)((const char *)base + ip * 8u);
}
__device__ __forceinline__ void stg_o2(double *base, unsigned ip, D2 v)
{
  *(__attribute__((address_space(1))) D2 *)((char *)base + ip * 8u) = v;
}
// the right-hand side pair of a lane from the FP64 vector, multiplied by
// dsc first (the value a scaling pass of the vector by dsc would have
// stored; dsc = 1 reads it as it is)
template <typename T, bool A16>
__device__ __forceinline__ typename V2T<T>::type ld_din(const double *din, unsigned ip,
                                                        double dsc)
{
  typedef typename V2T<T>::type T2;
  if (A16) {
    const D2 v = ldg_o2(din, ip);
    return T2{(T)(dsc * v.x), (T)(dsc * v.y)};
  }
  return T2{(T)(dsc * ldg_o(din, ip)), (T)(dsc * ldg_o(din, ip + 1))};
}

// a double kept in a slot of the interface address table
__device__ __forceinline__ double io_double(void *const *io, int q)
{
  double v;
  __builtin_memcpy(&v, &io[q], sizeof(double));
  return v;
}

template <typename T> struct PairSet {
  typedef typename V2T<T>::type T2;
  T2 c, bb, ae, an, as, au, xs, xn;
};

template <typename T, bool INIT, bool DIN = false, bool A16 = false>
__device__ __forceinline__ void load_pair(PairSet<T> &S, const Mat<T> &A,
                                          const T *__restrict__ xin, const T *__restrict__ b,
                                          unsigned ip, unsigned sy,
                                          const double *__restrict__ din = nullptr,
                                          double dsc = 1.0)
{
  typedef typename V2T<T>::type T2;
  S.c = ldo<T2>(A.c, ip);
  if (DIN) S.bb = ld_din<T, A16>(din, ip, dsc);
  else S.bb = ldo<T2>(b, ip);
  S.ae = ldo<T2>(A.u[0], ip);
  S.an = ldo<T2>(A.u[1], ip);
  S.as = ldo<T2>(A.u[1], ip - sy);
  S.au = ldo<T2>(A.u[2], ip);
  if (INIT) {
    S.xs = T2{T(0), T(0)};
    S.xn = T2{T(0), T(0)};
  } else {
    S.xs = ldo<T2>(xin, ip - sy);
    S.xn = ldo<T2>(xin, ip + sy);
  }
}

template <typename T, bool INIT>
__device__ __forceinline__ typename V2T<T>::type load_x2(const T *__restrict__ xin, unsigned ip)
{
  typedef typename V2T<T>::type T2;
  if (INIT) return T2{T(0), T(0)};
  return ldo<T2>(xin, ip);
}

// the YL form's plane loads: centre, b, east and upper couplings always;
// the south coupling and x_in of the row below only for the tile's lowest
// row (own_s), x_in of the row above only for its highest (own_n); the
// north coupling is loaded two planes ahead by the caller
template <typename T, bool INIT, bool DIN = false, bool A16 = false>
__device__ __forceinline__ void load_pair_yl(PairSet<T> &S, const Mat<T> &A,
                                             const T *__restrict__ xin,
                                             const T *__restrict__ b, unsigned ip, unsigned sy,
                                             bool own_s, bool own_n,
                                             const double *__restrict__ din = nullptr,
                                             double dsc = 1.0)
{
  typedef typename V2T<T>::type T2;
  S.c = ldo<T2>(A.c, ip);
  if (DIN) S.bb = ld_din<T, A16>(din, ip, dsc);
  else S.bb = ldo<T2>(b, ip);
  S.ae = ldo<T2>(A.u[0], ip);
  S.au = ldo<T2>(A.u[2], ip);
  S.xs = T2{T(0), T(0)};
  S.xn = T2{T(0), T(0)};
  S.as = T2{T(0), T(0)};
  if (own_s) {
    S.as = ldo<T2>(A.u[1], ip - sy);
    if (!INIT) S.xs = ldo<T2>(xin, ip - sy);
  }
  if (own_n && !INIT) S.xn = ldo<T2>(xin, ip + sy);
}

// FP32: ask for 8 waves per SIMD (<= 64 VGPRs), so two 1024-thread tiles fit
// on a CU; FP64 keeps its register budget
#define PAIR2_MIN_WAVES(T) (sizeof(T) == 4 ? 8 : 1)
// the YL form carries the north couplings two planes ahead and its outer
// rows' neighbour loads: up to 80 VGPRs (6 waves per SIMD) in FP32
#define PAIR2_YL_MIN_WAVES(T) (sizeof(T) == 4 ? 6 : 1)
// loading PF > 1 planes ahead keeps PF planes of loads in flight per wave
// across the per-plane barrier, at up to 128 VGPRs (4 waves per SIMD, the
// occupancy a 14-row tile of 16 waves has anyway)
#define PAIR2_PF_MIN_WAVES(T) (sizeof(T) == 4 ? 4 : 1)

// DOUT: the pair writes its result as FP64 into dout (a vector of the
// level's ghosted layout, no alignment assumed) instead of xout: the last
// sweep of a V-cycle hands the preconditioned vector straight to the FP64
// solver, without a conversion pass.  DIN (initial pairs): the right-hand
// side is read from the FP64 vector din and stored, converted, into bout
// (the level's b) at the pair's own points, which replaces the conversion
// pass at the start of the V-cycle.
//
// YL: the y neighbours' values (the south coupling, which is the row
// below's north one, and x_in of the rows below and above) come from LDS
// instead of three more loads per plane: every row publishes its north
// couplings and x_in of plane k+1 one plane ahead (both loaded two planes
// ahead), in an LDS buffer per plane parity, before the plane's barrier.
// Only the tile's outer rows load their outer neighbours.  6 loads per
// plane instead of 9, with the same single barrier per plane.
//
// PF (YL only): the coefficients are loaded PF planes ahead (x_in and the
// north couplings PF + 1), so PF planes of loads are in flight while a
// plane is updated; the arithmetic and its order do not depend on PF.
//
// A16 (DIN or DOUT): the FP64 vector's even-i points are 16-byte aligned,
// so a lane reads or writes its two values with one 16-byte access.
//
// DQ (with DOUT): the sums the difference-quotient J.v of the result z
// needs, u.z, z.z and sum |z| over the level's points (u: dqu), as per-
// workgroup partials dqp[q * nblocks + block] (k_dq_final adds them up in
// block order); u of plane k is loaded a plane before z of plane k is
// stored.
template <typename T, int TY, int KC, bool INIT, bool DOUT = false, bool DIN = false,
          bool YL = false, int PF = 1, bool A16 = false, bool DQ = false>
__global__ __launch_bounds__(64 * (TY + 2), (YL && PF > 1) || DQ ? PAIR2_PF_MIN_WAVES(T)
                                           : YL ? PAIR2_YL_MIN_WAVES(T) : PAIR2_MIN_WAVES(T))
void k_rbgs_pair2(LGeom g, Mat<T> A,
                                                              const T *__restrict__ xin,
                                                              T *__restrict__ xout,
                                                              const T *__restrict__ b,
                                                              double *__restrict__ dout,
                                                              const double *__restrict__ din,
                                                              T *__restrict__ bout,
                                                              void *const *io, double dsc,
                                                              const double *__restrict__ dqu,
                                                              double *__restrict__ dqp)
{
  typedef typename V2T<T>::type T2;
  static_assert(PF >= 1 && (YL || PF == 1), "prefetch depth > 1 needs the YL form");
  static_assert(!DQ || DOUT, "the J.v sums come with the FP64 result");
  // io: the FP64 interface vectors' addresses (and the right-hand side's
  // factor dsc) are read from device memory (a launch sequence captured
  // once and replayed with other vectors)
  if (DIN && io) {
    din = (const double *)io[0];
    dsc = io_double(io, 2);
  }
  if (DOUT && io) dout = (double *)io[1];
  if (DQ && io) dqu = (const double *)io[3];
  __shared__ T red[3][TY + 2][64];
  __shared__ double dqs[DQ ? 3 * (TY + 2) : 1];
  __shared__ T2 syn[YL ? 2 : 1][YL ? TY + 2 : 1][64];   // north couplings by plane parity
  __shared__ T2 syx[YL && !INIT ? 2 : 1][YL && !INIT ? TY + 2 : 1][64];   // x_in by parity
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = (g.n[0] + PX2 - 1) / PX2, gy = (g.n[1] + TY - 1) / TY;
  const int gz = (g.n[2] + KC - 1) / KC;
  const int nb = gx * gy * gz, nb8 = (nb + 7) & ~7;
  const int bid = blockIdx.x;
  const int Lt = (bid & 7) * (nb8 >> 3) + (bid >> 3);
  if (Lt >= nb) {                             // uniform over the workgroup
    if (DQ && tx == 0 && ty == 0) dqp[bid] = dqp[nb8 + bid] = dqp[2 * nb8 + bid] = 0.0;
    return;
  }
  const int bx = Lt % gx, rr = Lt / gx;
  const int by = rr % gy, bz = rr / gy;
  const int i0 = bx * PX2 - 2 + 2 * tx;       // even; the lane's points are i0, i0+1
  const int j = by * TY - 1 + ty;
  const bool valid = j >= 0 && j < g.n[1] && i0 >= 0 && i0 < g.n[0];
  const bool owned = valid && tx >= 1 && tx <= 62 && ty >= 1 && ty <= TY;
  const int nz = g.n[2];
  const int kb = bz * KC, ke = min(kb + KC, nz);
  const unsigned sy = g.s[1], sz = g.s[2];
  const bool hw0 = i0 > 0, he1 = i0 + 2 < g.n[0], hs = j > 0, hn = j + 1 < g.n[1];
  const int ic = min(max(i0, 0), g.n[0] - 2), jc = min(max(j, 0), g.n[1] - 1);
  const unsigned col = (unsigned)LIDX(g, ic, jc, 0) - sz;   // plane q at col + (q+1) sz
  auto at = [&](int q) { return col + (unsigned)(min(max(q, 0), nz - 1) + 1) * sz; };
  // the march starts one plane below the chunk (its red values feed the
  // chunk's first black update), except at the bottom of the level, where
  // that plane does not exist
  const int k0 = (kb == 0) ? 0 : kb - 1;
  unsigned ip = valid ? (unsigned)LIDX(g, i0, j, k0) : 0u;   // store address, plane k-1
  // YL: the tile's outer rows still load their outer neighbours (uniform
  // over each row's wavefront)
  const bool own_s = !YL || ty == 0, own_n = !YL || ty == TY + 1;

  // carried state
  T sc = T(1), sb = T(0), saw = T(0), sae = T(0), sas = T(0), san = T(0), sal = T(0), sau = T(0);
  T rm1 = T(0), rm2 = T(0);
  // DQ: u of planes k (uq) and k-1 (up), the lane's running sums
  double uq0 = 0.0, uq1 = 0.0, up0 = 0.0, up1 = 0.0;
  double q_uz = 0.0, q_zz = 0.0, q_z1 = 0.0;
  // at the top of the step for plane k: cur = plane k, ahead[d] = plane
  // k+1+d (d < PF-1), anr[d] = north couplings of plane k+1+d (d < PF),
  // xr[d] = x_in of plane k-1+d (d < PF+2); the step loads ahead[PF-1],
  // anr[PF] and xr[PF+2]
  PairSet<T> cur, ahead[PF];
  T2 anr[PF + 1];
  T2 xr[PF + 3];
#pragma unroll
  for (int d = 0; d <= PF; d++) anr[d] = T2{T(0), T(0)};
  if (YL) {
    load_pair_yl<T, INIT, DIN, A16>(cur, A, xin, b, at(k0), sy, own_s, own_n, din, dsc);
    cur.an = ldo<T2>(A.u[1], at(k0));
#pragma unroll
    for (int d = 0; d + 1 < PF; d++)
      load_pair_yl<T, INIT, DIN, A16>(ahead[d], A, xin, b, at(k0 + 1 + d), sy, own_s, own_n,
                                      din, dsc);
#pragma unroll
    for (int d = 0; d < PF; d++) anr[d] = ldo<T2>(A.u[1], at(k0 + 1 + d));
  } else {
    load_pair<T, INIT, DIN, A16>(cur, A, xin, b, at(k0), sy, din, dsc);
  }
  T2 al = ldo<T2>(A.u[2], at(k0 - 1));               // lower coupling of plane k0
#pragma unroll
  for (int d = 0; d < PF + 2; d++) xr[d] = load_x2<T, INIT>(xin, at(k0 - 1 + d));
  if (YL) {
    // plane k0's shared values, before the first plane's red update
    syn[k0 & 1][ty][tx] = cur.an;
    if (!INIT) syx[k0 & 1][ty][tx] = xr[1];
    lds_barrier();
  }
  int it = 0;
  for (int k = k0; k <= ke; k++, it++) {
    // prefetch: coefficients and y-neighbour x of plane k+PF, x of plane
    // k+PF+1 (YL: the north couplings of plane k+PF+1 as well, published
    // one plane before their own)
    if (YL) {
      load_pair_yl<T, INIT, DIN, A16>(ahead[PF - 1], A, xin, b, at(k + PF), sy, own_s, own_n,
                                      din, dsc);
      anr[PF] = ldo<T2>(A.u[1], at(k + PF + 1));
    } else {
      load_pair<T, INIT, DIN, A16>(ahead[0], A, xin, b, at(k + 1), sy, din, dsc);
    }
    xr[PF + 2] = load_x2<T, INIT>(xin, at(k + PF + 1));
    if (DQ) {
      // u of plane k, for the sums when z of plane k is stored (next step)
      uq0 = ldg_o(dqu, ip);
      uq1 = ldg_o(dqu, ip + 1);
    }
    const T2 xm = xr[0], x0 = xr[1], xp = xr[2];
    if (DIN && owned && k >= kb && k < ke) sto<T2>(bout, ip, cur.bb);
    if (YL) {
      // plane k+1's north couplings and x_in for the rows beside
      syn[(k + 1) & 1][ty][tx] = anr[0];
      if (!INIT) syx[(k + 1) & 1][ty][tx] = xp;
      // this plane's y neighbours, published one plane ago
      if (!own_s) {
        cur.as = syn[k & 1][ty - 1][tx];
        if (!INIT) cur.xs = syx[k & 1][ty - 1][tx];
      }
      if (!own_n && !INIT) cur.xn = syx[k & 1][ty + 1][tx];
    }
    // ---- red at plane k: the lane's red point is 0 when (i0+j+k) is odd
    const bool p0red = ((i0 + j + k) & 1) == 1;
    const bool here = valid && k >= 0 && k < nz;
    const T xw0 = from_prev(x0.y), xe1 = from_next(x0.x);
    const T aw0 = from_prev(cur.ae.y);
    T rr;
    {
      T c = p0red ? cur.c.x : cur.c.y, bb = p0red ? cur.bb.x : cur.bb.y;
      if (INIT) {
        rr = gs_div(bb, c);
      } else {
        T aw = p0red ? aw0 : cur.ae.x, xw = p0red ? xw0 : x0.x;
        bool hw = p0red ? hw0 : true;
        T ae = p0red ? cur.ae.x : cur.ae.y, xe = p0red ? x0.y : xe1;
        bool he = p0red ? true : he1;
        T as = p0red ? cur.as.x : cur.as.y, xs = p0red ? cur.xs.x : cur.xs.y;
        T an = p0red ? cur.an.x : cur.an.y, xn = p0red ? cur.xn.x : cur.xn.y;
        T alv = p0red ? al.x : al.y, xl = p0red ? xm.x : xm.y;
        T au = p0red ? cur.au.x : cur.au.y, xu = p0red ? xp.x : xp.y;
        rr = gs_update(bb, c, aw, xw, hw, ae, xe, he, as, xs, hs, an, xn, hn, alv, xl, k > 0,
                       au, xu, k + 1 < nz);
      }
    }
    const T rk = here ? rr : T(0);
    red[it % 3][ty][tx] = rk;
    // ---- black at plane k-1: the lane's black point there is its red
    // point of plane k; x neighbours are red values of plane k-1 (rm1 is
    // the lane's own, the lane beside supplies the other)
    const int kk = k - 1;
    const T rprev = from_prev(rm1), rnext = from_next(rm1);
    {
      const T(*R)[64] = red[(it + 2) % 3];
      const int tym = ty > 0 ? ty - 1 : 0, typ = ty < TY + 1 ? ty + 1 : TY + 1;
      T xw = p0red ? rprev : rm1, xe = p0red ? rm1 : rnext;
      bool hw = p0red ? hw0 : true, he = p0red ? true : he1;
      T vb = gs_update(sb, sc, saw, xw, hw, sae, xe, he, sas, R[tym][tx], hs, san, R[typ][tx], hn,
                       sal, rm2, kk > 0, sau, rk, kk + 1 < nz);
      if (owned && kk >= kb && kk < ke) {
        T2 out = p0red ? T2{vb, rm1} : T2{rm1, vb};
        if (DOUT && A16) {
          stg_o2(dout, ip - sz, D2{(double)out.x, (double)out.y});
        } else if (DOUT) {
          stg_o(dout, ip - sz, (double)out.x);
          stg_o(dout, ip - sz + 1, (double)out.y);
        } else {
          sto<T2>(xout, ip - sz, out);
        }
        if (DQ) {
          const double z0 = (double)out.x, z1 = (double)out.y;
          q_uz += up0 * z0;
          q_uz += up1 * z1;
          q_zz += z0 * z0;
          q_zz += z1 * z1;
          q_z1 += fabs(z0);
          q_z1 += fabs(z1);
        }
      }
    }
    // the black point of plane k keeps its coefficients for the next step
    // (after the black update of plane k-1, which still needs the old set)
    if (here) {
      sc = p0red ? cur.c.y : cur.c.x;   sb = p0red ? cur.bb.y : cur.bb.x;
      saw = p0red ? cur.ae.x : aw0;     sae = p0red ? cur.ae.y : cur.ae.x;
      sas = p0red ? cur.as.y : cur.as.x; san = p0red ? cur.an.y : cur.an.x;
      sal = p0red ? al.y : al.x;        sau = p0red ? cur.au.y : cur.au.x;
    }
    lds_barrier();
    rm2 = rm1;
    rm1 = rk;
    al = cur.au;
    cur = ahead[0];
#pragma unroll
    for (int d = 0; d + 1 < PF; d++) ahead[d] = ahead[d + 1];
    if (YL) {
      cur.an = anr[0];
#pragma unroll
      for (int d = 0; d < PF; d++) anr[d] = anr[d + 1];
    }
#pragma unroll
    for (int d = 0; d < PF + 2; d++) xr[d] = xr[d + 1];
    if (DQ) { up0 = uq0; up1 = uq1; }
    ip += sz;
  }
  if (DQ) {
    // per wavefront (one tile row), then the rows in order
    const double w_uz = wave_sum_d(q_uz), w_zz = wave_sum_d(q_zz), w_z1 = wave_sum_d(q_z1);
    if (tx == 0) {
      dqs[3 * ty] = w_uz;
      dqs[3 * ty + 1] = w_zz;
      dqs[3 * ty + 2] = w_z1;
    }
    __syncthreads();
    if (tx == 0 && ty == 0) {
      double a = 0.0, bq = 0.0, c = 0.0;
      for (int r = 0; r < TY + 2; r++) {
        a += dqs[3 * r];
        bq += dqs[3 * r + 1];
        c += dqs[3 * r + 2];
      }
      dqp[bid] = a;
      dqp[nb8 + bid] = bq;
      dqp[2 * nb8 + bid] = c;
    }
  }
}

#define DQ_ROW_BLOCKS 1024
// the J.v sums u.z, z.z and sum |z| over the level's points as per-block
// partials dqp[q * gridDim.x + block], for V-cycles whose last pair does
// not write the FP64 result (io: z and u from the interface table)
__global__ __launch_bounds__(256) void k_dq_rows(LGeom g, const double *__restrict__ u,
                                                 const double *__restrict__ z,
                                                 double *__restrict__ dqp, void *const *io)
{
  if (io) {
    z = (const double *)io[1];
    u = (const double *)io[3];
  }
  __shared__ double part[3][256];
  double a = 0.0, b = 0.0, c = 0.0;
  const int rows = g.n[1] * g.n[2];
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const int j = r % g.n[1], k = r / g.n[1];
    for (int i = threadIdx.x; i < g.n[0]; i += blockDim.x) {
      const long ip = LIDX(g, i, j, k);
      const double zv = z[ip], uv = u[ip];
      a += uv * zv;
      b += zv * zv;
      c += fabs(zv);
    }
  }
  part[0][threadIdx.x] = a;
  part[1][threadIdx.x] = b;
  part[2][threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 3; q++) part[q][threadIdx.x] += part[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; q++) dqp[q * gridDim.x + blockIdx.x] = part[q][0];
}

// the J.v sums of k_rbgs_pair2<..., DQ> (or k_dq_rows): dst[q] = sum over
// the nb8 block partials dqp[q * nb8 + block], in block order (one
// workgroup)
__global__ __launch_bounds__(256) void k_dq_final(const double *__restrict__ dqp, int nb8,
                                                  double *__restrict__ dst)
{
  __shared__ double part[256];
  for (int q = 0; q < 3; q++) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nb8; i += 256) a += dqp[q * nb8 + i];
    part[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) dst[q] = part[0];
    __syncthreads();
  }
}


// ---- level residual fused with the z restriction ---------------------------
// For a level coarsened in z with an even x extent: r = b - A x on the fine
// planes of a z-chunk (resid_point's terms and order), restricted to the
// coarse planes as the chunk marches up (restrict_point's weights and
// order), so r is never stored and re-read.  Lanes own two x-adjacent points
// as in k_rbgs_pair2 (lanes 0 and 63 supply the x neighbours at the tile
// edges); a row of the tile is one wavefront; the z neighbours and the lower
// coupling are carried from plane to plane.  A chunk of KC (even) fine
// planes produces the coarse planes centred on its even planes and
// recomputes the fine plane below it for the first one's lower weight.
template <typename T> struct RRSet {
  typedef typename V2T<T>::type T2;
  T2 c, bb, ae, an, as, au, xs, xn;
};

template <typename T>
__device__ __forceinline__ void load_rr(RRSet<T> &S, const Mat<T> &A, const T *__restrict__ x,
                                        const T *__restrict__ b, unsigned ip, unsigned sy)
{
  typedef typename V2T<T>::type T2;
  S.c = *(const T2 *)(A.c + ip);
  S.bb = *(const T2 *)(b + ip);
  S.ae = *(const T2 *)(A.u[0] + ip);
  S.an = *(const T2 *)(A.u[1] + ip);
  S.as = *(const T2 *)(A.u[1] + ip - sy);
  S.au = *(const T2 *)(A.u[2] + ip);
  S.xs = *(const T2 *)(x + ip - sy);
  S.xn = *(const T2 *)(x + ip + sy);
}

template <typename T, int TY, int KC>
__global__ __launch_bounds__(64 * TY) void k_resid_restrict_z(LGeom gf, LGeom gc, Mat<T> A,
                                                              const T *__restrict__ x,
                                                              const T *__restrict__ b,
                                                              T *__restrict__ rc,
                                                              const T *__restrict__ p1,
                                                              const T *__restrict__ p2)
{
  typedef typename V2T<T>::type T2;
  static_assert(KC % 2 == 0, "z-chunks start on coarse points");
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = (gf.n[0] + PX2 - 1) / PX2, gy = (gf.n[1] + TY - 1) / TY;
  const int gz = (gf.n[2] + KC - 1) / KC;
  const int nb = gx * gy * gz, nb8 = (nb + 7) & ~7;
  const int bid = blockIdx.x;
  const int Lt = (bid & 7) * (nb8 >> 3) + (bid >> 3);
  if (Lt >= nb) return;                       // uniform over the workgroup
  const int bx = Lt % gx, rr = Lt / gx;
  const int by = rr % gy, bz = rr / gy;
  const int i0 = bx * PX2 - 2 + 2 * tx;       // even; the lane's points are i0, i0+1
  const int j = by * TY + ty;
  const bool valid = j < gf.n[1] && i0 >= 0 && i0 < gf.n[0];
  const bool owned = valid && tx >= 1 && tx <= 62;
  const int nz = gf.n[2];
  const int kb = bz * KC, ke = min(kb + KC, nz);
  const unsigned sy = gf.s[1], sz = gf.s[2];
  const bool hw0 = i0 > 0, he1 = i0 + 2 < gf.n[0], hs = j > 0, hn = j + 1 < gf.n[1];
  // loads come from the clamped column and planes in [0, nz-1]
  const int ic = min(max(i0, 0), gf.n[0] - 2), jc = min(j, gf.n[1] - 1);
  const unsigned col = (unsigned)LIDX(gf, ic, jc, 0) - sz;   // plane q at col + (q+1) sz
  auto at = [&](int q) { return col + (unsigned)(min(max(q, 0), nz - 1) + 1) * sz; };
  // coarse store address (same x-y layout as the fine level), coarse plane kb/2
  unsigned ipc = valid ? (unsigned)LIDX(gc, i0, j, kb / 2) : 0u;
  const unsigned szc = gc.s[2];

  const int k0 = (kb > 0) ? kb - 1 : 0;
  RRSet<T> cur, nxt;
  load_rr<T>(cur, A, x, b, at(k0), sy);
  T2 al = *(const T2 *)(A.u[2] + at(k0 - 1));         // lower coupling of plane k0
  T2 xm = *(const T2 *)(x + at(k0 - 1));
  T2 x0 = *(const T2 *)(x + at(k0));
  T2 xp = *(const T2 *)(x + at(k0 + 1));
  T2 rF = T2{T(0), T(0)}, lo = T2{T(0), T(0)};
  for (int k = k0; k < ke; k++) {
    load_rr<T>(nxt, A, x, b, at(k + 1), sy);
    const T2 xpp = *(const T2 *)(x + at(k + 2));
    const bool odd = (k & 1) != 0;
    T2 w1 = T2{T(0), T(0)}, w2 = T2{T(0), T(0)};
    if (odd) {                                          // uniform over the workgroup
      w1 = *(const T2 *)(p1 + at(k));
      w2 = *(const T2 *)(p2 + at(k));
    }
    // r at plane k, both points
    const T xw0 = from_prev(x0.y), xe1 = from_next(x0.x);
    const T aw0 = from_prev(cur.ae.y);
    T v0 = cur.c.x * x0.x;
    if (hw0) v0 += aw0 * xw0;
    v0 += cur.ae.x * x0.y;
    if (hs) v0 += cur.as.x * cur.xs.x;
    if (hn) v0 += cur.an.x * cur.xn.x;
    if (k > 0) v0 += al.x * xm.x;
    if (k + 1 < nz) v0 += cur.au.x * xp.x;
    T v1 = cur.c.y * x0.y;
    v1 += cur.ae.x * x0.x;
    if (he1) v1 += cur.ae.y * xe1;
    if (hs) v1 += cur.as.y * cur.xs.y;
    if (hn) v1 += cur.an.y * cur.xn.y;
    if (k > 0) v1 += al.y * xm.y;
    if (k + 1 < nz) v1 += cur.au.y * xp.y;
    const T2 r = T2{cur.bb.x - v0, cur.bb.y - v1};
    if (!odd) {
      // a coarse centre: its lower weight came from plane k-1 (none at k = 0)
      rF = r;
      if (k == 0) lo = T2{T(0), T(0)};
      if (k + 1 >= nz) {                               // no plane above: hi = 0
        if (owned) *(T2 *)(rc + ipc) = T2{rF.x + (lo.x + T(0)), rF.y + (lo.y + T(0))};
        ipc += szc;
      }
    } else {
      if (k > kb) {                                     // finish the centre at k-1
        const T2 hi = T2{w1.x * r.x, w1.y * r.y};
        if (owned) *(T2 *)(rc + ipc) = T2{rF.x + (lo.x + hi.x), rF.y + (lo.y + hi.y)};
        ipc += szc;
      }
      lo = T2{w2.x * r.x, w2.y * r.y};                 // lower weight of the centre at k+1
    }
    al = cur.au;
    cur = nxt;
    xm = x0; x0 = xp; xp = xpp;
  }
}

// level residual and restriction in one pass, one thread per coarse point:
// the fine residuals at C-1, C, C+1 along the coarsening direction d
// (resid_point's terms and order; a fine point between two coarse ones is
// evaluated by both) combined with restrict_point's weights and order, so
// the residual is never stored and each small level costs one launch, not two
template <typename T, bool FULL>
__global__ __launch_bounds__(256) void k_resid_restrict(LGeom gf, LGeom gc, int d, Mat<T> A,
                                                        const T *__restrict__ x,
                                                        const T *__restrict__ b,
                                                        T *__restrict__ rc,
                                                        const T *__restrict__ p1,
                                                        const T *__restrict__ p2)
{
  XCD_IJK(gc);
  int f[3] = { i, j, k };
  f[d] *= 2;
  const long iF = LIDX(gf, f[0], f[1], f[2]);
  const long sd = gf.s[d];
  const T rC = resid_value<T, FULL>(gf, A, x, b, f[0], f[1], f[2]);
  T lo = T(0), hi = T(0);
  if (f[d] > 0) {
    int g[3] = { f[0], f[1], f[2] };
    g[d] -= 1;
    lo = p2[iF - sd] * resid_value<T, FULL>(gf, A, x, b, g[0], g[1], g[2]);
  }
  if (f[d] + 1 < gf.n[d]) {
    int g[3] = { f[0], f[1], f[2] };
    g[d] += 1;
    hi = p1[iF + sd] * resid_value<T, FULL>(gf, A, x, b, g[0], g[1], g[2]);
  }
  rc[LIDX(gc, i, j, k)] = rC + (lo + hi);
}

template <typename T>
__global__ __launch_bounds__(256) void k_prolong_add(LGeom gf, LGeom gc, int d,
                                                     const T *__restrict__ ec,
                                                     T *__restrict__ xf,
                                                     const T *__restrict__ p1,
                                                     const T *__restrict__ p2)
{
  XCD_IJK(gf);
  prolong_point(gf, gc, d, ec, xf, p1, p2, i, j, k);
}

// x_f += P e_c on a z-coarsened level as a z-march over (i, j) columns (64 x
// 4 column tiles, XCD-aware): each coarse value is loaded once and used by
// the even plane it is injected into and the two odd planes beside it; the
// next planes' loads are issued before the current ones are added.  Same
// expressions as prolong_point.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_z(LGeom gf, LGeom gc, const T *__restrict__ ec,
                                                   T *__restrict__ xf,
                                                   const T *__restrict__ p1,
                                                   const T *__restrict__ p2)
{
  const int gx = (gf.n[0] + 63) >> 6, gy = (gf.n[1] + 3) >> 2;
  const int nb = gx * gy, nb8 = (nb + 7) & ~7;
  const int b = blockIdx.x;
  const int L = (b & 7) * (nb8 >> 3) + (b >> 3);
  if (L >= nb) return;
  const int i = (L % gx) * 64 + threadIdx.x, j = (L / gx) * 4 + threadIdx.y;
  if (i >= gf.n[0] || j >= gf.n[1]) return;
  const long sf = gf.s[2], sc = gc.s[2];
  const int nfz = gf.n[2], ncz = gc.n[2];
  long ip = LIDX(gf, i, j, 0);
  long ic = LIDX(gc, i, j, 0);
  // plane pair (2kc, 2kc+1): coarse values kc and kc+1 (the ghost plane,
  // zero, above the last), fine x of both planes, weights of the odd one
  T e0 = ec[ic], e1 = ec[ic + sc];
  T xe = xf[ip], xo = xf[ip + sf], w1 = p1[ip + sf], w2 = p2[ip + sf];
  for (int kc = 0; kc < ncz; kc++) {
    const int fk = 2 * kc;
    T ne1 = T(0), nxe = T(0), nxo = T(0), nw1 = T(0), nw2 = T(0);
    if (kc + 1 < ncz) {
      ne1 = ec[ic + 2 * sc];               // at most the ghost plane above
      nxe = xf[ip + 2 * sf];
      nxo = xf[ip + 3 * sf];               // at most the ghost plane above
      nw1 = p1[ip + 3 * sf];
      nw2 = p2[ip + 3 * sf];
    }
    xf[ip] = xe + e0;
    if (fk + 1 < nfz) {
      T lo = w1 * e0;
      T hi = (fk + 2 < nfz) ? w2 * e1 : T(0);
      xf[ip + sf] = xo + (lo + hi);
    }
    e0 = e1; e1 = ne1;
    xe = nxe; xo = nxo; w1 = nw1; w2 = nw2;
    ip += 2 * sf;
    ic += sc;
  }
}

// k_convert with the FP64 interface vector's address read from io[q] (q = 0:
// the source is the right-hand side, q = 1: the destination is the result),
// for launch sequences captured once and replayed with other vectors
// (q = 0 multiplies the source by the right-hand side's factor in io[2]
// first, as k_convert_in does)
template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert_io(LGeom g, const S *__restrict__ src,
                                                    D *__restrict__ dst, void *const *io, int q)
{
  XCD_IJK(g);
  const S *sp = (q == 0) ? (const S *)io[0] : src;
  D *dp = (q == 1) ? (D *)io[1] : dst;
  long ip = LIDX(g, i, j, k);
  dp[ip] = (q == 0) ? (D)(io_double(io, 2) * (double)sp[ip]) : (D)sp[ip];
}

// the interface addresses of one replay of a captured V-cycle, and the
// right-hand side's factor
__global__ void k_set_io(void **io, const double *b, double *x, double bscale,
                         const double *u)
{
  io[0] = (void *)b;
  io[1] = (void *)x;
  __builtin_memcpy(&io[2], &bscale, sizeof(double));
  io[3] = (void *)u;
}

// the V-cycle's right-hand side in the level precision: sc * src, the
// value a scaling pass of src by sc would have stored, converted
template <typename T>
__global__ __launch_bounds__(256) void k_convert_in(LGeom g, const double *__restrict__ src,
                                                    T *__restrict__ dst, double sc)
{
  XCD_IJK(g);
  long ip = LIDX(g, i, j, k);
  dst[ip] = (T)(sc * src[ip]);
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert(LGeom g, const S *__restrict__ src,
                                                 D *__restrict__ dst)
{
  XCD_IJK(g);
  long ip = LIDX(g, i, j, k);
  dst[ip] = (D)src[ip];
}

template <typename T>
__global__ void k_fill_t(T *p, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (long)gridDim.x * blockDim.x) p[i] = T(0);
}

// ---- single-workgroup kernels for the small levels ------------------------
// Phases over a level's points; every phase ends with a block barrier.
#define FOR_POINTS(g, body)                                              \
  {                                                                      \
    int np_ = (g).n[0] * (g).n[1] * (g).n[2];                            \
    for (int idx_ = threadIdx.x; idx_ < np_; idx_ += blockDim.x) {       \
      int i, j, k;                                                       \
      unflat((g), idx_, i, j, k);                                        \
      body;                                                              \
    }                                                                    \
    __syncthreads();                                                     \
  }

// One half-sweep of parity par over a level inside a workgroup.  Threads
// enumerate the points of that parity only ((i + j + k) & 1 == par), so no
// lane idles on the parity test; WUNR points per thread per pass have their
// loads issued before their updates (1 measured fastest at 1024 threads:
// 2 and 4 raise register use past the 1024-thread budget).  The arithmetic
// per point is rbgs_point's.
#define WUNR 1
template <typename T, bool FULL>
__device__ void rbgs_phase(const LGeom &g, const Mat<T> &A, T *__restrict__ x,
                           const T *__restrict__ b, int par, int init)
{
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int ncol = (nx + 1) >> 1;
  const int np = ncol * ny * nz;
  const int sy = g.s[1], sz = g.s[2];
  for (int base = threadIdx.x; base < np; base += WUNR * blockDim.x) {
    T bb[WUNR], c[WUNR], aw[WUNR], xw[WUNR], ae[WUNR], xe[WUNR], as[WUNR], xs[WUNR];
    T an[WUNR], xn[WUNR], al[WUNR], xl[WUNR], au[WUNR], xu[WUNR];
    bool hw[WUNR], he[WUNR], hs[WUNR], hn[WUNR], hl[WUNR], hu[WUNR], act[WUNR];
    int ip[WUNR];
#pragma unroll
    for (int u = 0; u < WUNR; u++) {
      const int idx = base + u * blockDim.x;
      const int t = idx % ncol, r = idx / ncol;
      const int j = r % ny, k = r / ny;
      const int i = ((par - j - k) & 1) + 2 * t;
      const bool a = (idx < np) && (i < nx);
      act[u] = a;
      ip[u] = a ? LIDX(g, i, j, k) : 0;
      const int q = ip[u];
      hw[u] = a && i > 0; he[u] = a && i + 1 < nx; hs[u] = a && j > 0;
      hn[u] = a && j + 1 < ny; hl[u] = a && k > 0; hu[u] = a && k + 1 < nz;
      bb[u] = a ? b[q] : T(0);
      c[u] = a ? A.c[q] : T(1);
      const bool nb = a && !init;
      aw[u] = (nb && hw[u]) ? lo_cpl<T, FULL>(A, q, 0, 1) : T(0);
      xw[u] = (nb && hw[u]) ? x[q - 1] : T(0);
      ae[u] = nb ? A.u[0][q] : T(0);
      xe[u] = (nb && he[u]) ? x[q + 1] : T(0);
      as[u] = (nb && hs[u]) ? lo_cpl<T, FULL>(A, q, 1, sy) : T(0);
      xs[u] = (nb && hs[u]) ? x[q - sy] : T(0);
      an[u] = nb ? A.u[1][q] : T(0);
      xn[u] = (nb && hn[u]) ? x[q + sy] : T(0);
      al[u] = (nb && hl[u]) ? lo_cpl<T, FULL>(A, q, 2, sz) : T(0);
      xl[u] = (nb && hl[u]) ? x[q - sz] : T(0);
      au[u] = nb ? A.u[2][q] : T(0);
      xu[u] = (nb && hu[u]) ? x[q + sz] : T(0);
    }
#pragma unroll
    for (int u = 0; u < WUNR; u++) {
      if (!act[u]) continue;
      x[ip[u]] = init ? gs_div(bb[u], c[u])
                      : gs_update(bb[u], c[u], aw[u], xw[u], hw[u], ae[u], xe[u], he[u], as[u],
                                  xs[u], hs[u], an[u], xn[u], hn[u], al[u], xl[u], hl[u], au[u],
                                  xu[u], hu[u]);
    }
  }
  __syncthreads();
}

// t = b - A x over a level inside a workgroup, WUNR points per thread per
// pass with the loads first; resid_point's expressions and order
template <typename T, bool FULL>
__device__ void resid_phase(const LGeom &g, const Mat<T> &A, const T *__restrict__ x,
                            const T *__restrict__ b, T *__restrict__ tt)
{
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int np = nx * ny * nz;
  const int sy = g.s[1], sz = g.s[2];
  for (int base = threadIdx.x; base < np; base += WUNR * blockDim.x) {
    T bb[WUNR], c[WUNR], x0[WUNR], aw[WUNR], xw[WUNR], ae[WUNR], xe[WUNR], as[WUNR], xs[WUNR];
    T an[WUNR], xn[WUNR], al[WUNR], xl[WUNR], au[WUNR], xu[WUNR];
    bool hw[WUNR], he[WUNR], hs[WUNR], hn[WUNR], hl[WUNR], hu[WUNR], act[WUNR];
    int ip[WUNR];
#pragma unroll
    for (int u = 0; u < WUNR; u++) {
      const int idx = base + u * blockDim.x;
      const bool a = idx < np;
      int i = 0, j = 0, k = 0;
      if (a) unflat(g, idx, i, j, k);
      act[u] = a;
      const int q = a ? LIDX(g, i, j, k) : 0;
      ip[u] = q;
      hw[u] = a && i > 0; he[u] = a && i + 1 < nx; hs[u] = a && j > 0;
      hn[u] = a && j + 1 < ny; hl[u] = a && k > 0; hu[u] = a && k + 1 < nz;
      bb[u] = a ? b[q] : T(0);
      c[u] = a ? A.c[q] : T(0);
      x0[u] = a ? x[q] : T(0);
      aw[u] = hw[u] ? lo_cpl<T, FULL>(A, q, 0, 1) : T(0);
      xw[u] = hw[u] ? x[q - 1] : T(0);
      ae[u] = he[u] ? A.u[0][q] : T(0);
      xe[u] = he[u] ? x[q + 1] : T(0);
      as[u] = hs[u] ? lo_cpl<T, FULL>(A, q, 1, sy) : T(0);
      xs[u] = hs[u] ? x[q - sy] : T(0);
      an[u] = hn[u] ? A.u[1][q] : T(0);
      xn[u] = hn[u] ? x[q + sy] : T(0);
      al[u] = hl[u] ? lo_cpl<T, FULL>(A, q, 2, sz) : T(0);
      xl[u] = hl[u] ? x[q - sz] : T(0);
      au[u] = hu[u] ? A.u[2][q] : T(0);
      xu[u] = hu[u] ? x[q + sz] : T(0);
    }
#pragma unroll
    for (int u = 0; u < WUNR; u++) {
      if (!act[u]) continue;
      T v = c[u] * x0[u];
      if (hw[u]) v += aw[u] * xw[u];
      if (he[u]) v += ae[u] * xe[u];
      if (hs[u]) v += as[u] * xs[u];
      if (hn[u]) v += an[u] * xn[u];
      if (hl[u]) v += al[u] * xl[u];
      if (hu[u]) v += au[u] * xu[u];
      tt[ip[u]] = bb[u] - v;
    }
  }
  __syncthreads();
}

// RedBlackGSPoint call sequence (PL/rb_GS_point.c:100-300) inside a block
template <typename T, bool FULL>
__device__ void rbgs_block(const LevelDev<T> &L, int zero, int max_iter, int symmetric)
{
  const LGeom &g = L.g;
  int count = symmetric ? 1 : 2;
  int rb = 0;
  if (zero) {
    count++;
    if ((count / 2) > max_iter) {
      FOR_POINTS(g, L.x[LIDX(g, i, j, k)] = T(0));
      if (!symmetric) return;
    }
    rbgs_phase<T, FULL>(g, L.A, L.x, L.b, 1, 1);
    rb = !rb;
  }
  for (; (count / 2) <= max_iter; count++) {
    int par = rb ? 0 : 1;
    rbgs_phase<T, FULL>(g, L.A, L.x, L.b, par, 0);
    rb = !rb;
  }
}

// V-cycle over levels lc .. L-1 inside one workgroup (b[lc] given, x[lc]
// returned)
template <typename T, bool FULL>
__device__ void vcycle_block(const LevelDev<T> *__restrict__ lev, int lc, int L, int pre,
                             int post, int symmetric)
{
  int l;
  for (l = lc; l <= L - 2; l++) {
    const LevelDev<T> &F = lev[l], &Cc = lev[l + 1];
    rbgs_block<T, FULL>(F, 1, pre, symmetric);
    resid_phase<T, FULL>(F.g, F.A, F.x, F.b, F.t);
    FOR_POINTS(Cc.g, restrict_point(F.g, Cc.g, F.dir, F.t, Cc.b, F.p1, F.p2, i, j, k));
  }
  rbgs_block<T, FULL>(lev[L - 1], 1, 1, symmetric);
  for (l = L - 2; l >= lc; l--) {
    const LevelDev<T> &F = lev[l], &Cc = lev[l + 1];
    FOR_POINTS(F.g, prolong_point(F.g, Cc.g, F.dir, Cc.x, F.x, F.p1, F.p2, i, j, k));
    rbgs_block<T, FULL>(F, 0, post, symmetric);
  }
}

// bottom of the V-cycle: levels lc .. L-1 (b[lc] given, x[lc] returned)
template <typename T, bool FULL>
__global__ __launch_bounds__(CWG) void k_coarse_vcycle(const LevelDev<T> *__restrict__ lev,
                                                       int lc, int L, int pre, int post,
                                                       int symmetric)
{
  vcycle_block<T, FULL>(lev, lc, L, pre, post, symmetric);
}

// The same bottom with every level's arrays in LDS (symmetric storage).
// Each level is copied into a compact, ghost-free block of LDS (C, E, N,
// U, x, b, t, p1, p2 of its points); its LDS level record carries compact
// strides and array pointers shifted so that LIDX addresses the compact
// block.  Every per-point function reads a neighbour or a lower coupling
// only when that neighbour exists, so no ghost entry is ever touched.  The
// phases then wait on LDS instead of L2, and the sweeps, residuals and
// transfers are the global kernel's (the same functions, in the same
// order).  b of the top level is copied in, x of the top level out; the
// coarser levels' x, b, t are produced inside.
#define LDS_ALIGN 16
__host__ __device__ static inline long lds_level_table_bytes(int nl, long rec)
{
  return ((nl * rec) + LDS_ALIGN - 1) & ~(long)(LDS_ALIGN - 1);
}

// bin / xout (the ensemble's levels 0 .. L-1 in one workgroup): the top
// level's b comes from the FP64 vector bin and its x goes to xout, both in
// the top level's ghosted layout; otherwise b and x are the top level's own
template <typename T>
__device__ void vcycle_lds(const LevelDev<T> *__restrict__ lev, int lc, int L, int pre,
                           int post, int symmetric, const double *__restrict__ bin,
                           double *__restrict__ xout)
{
  extern __shared__ __align__(LDS_ALIGN) unsigned char smem[];
  const int nl = L - lc;
  LevelDev<T> *lv = (LevelDev<T> *)smem;
  T *pool = (T *)(smem + lds_level_table_bytes(nl, (long)sizeof(LevelDev<T>)));
  if (threadIdx.x == 0) {
    T *q = pool;
    for (int m = 0; m < nl; m++) {
      const LevelDev<T> &G = lev[lc + m];
      LevelDev<T> D;
      D.g = G.g;
      D.g.s[0] = 1;
      D.g.s[1] = G.g.n[0];
      D.g.s[2] = G.g.n[0] * G.g.n[1];
      D.dir = G.dir; D.o1 = G.o1; D.o2 = G.o2;
      const long np = (long)G.g.n[0] * G.g.n[1] * G.g.n[2];
      const long sh = 1 + D.g.s[1] + D.g.s[2];   // LIDX of (0, 0, 0)
      D.A.c = q - sh; q += np;
      for (int d = 0; d < 3; d++) { D.A.u[d] = q - sh; q += np; D.A.l[d] = nullptr; }
      D.x = q - sh; q += np;
      D.b = q - sh; q += np;
      D.t = q - sh; q += np;
      D.p1 = q - sh; q += np;
      D.p2 = q - sh; q += np;
      lv[m] = D;
    }
  }
  __syncthreads();
  for (int m = 0; m < nl; m++) {
    const LevelDev<T> &G = lev[lc + m];
    const LevelDev<T> &D = lv[m];
    const bool top = (m == 0);
    const bool last = (m == nl - 1);
    const int np = G.g.n[0] * G.g.n[1] * G.g.n[2];
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) {
      int i, j, k;
      unflat(G.g, idx, i, j, k);
      const long ig = LIDX(G.g, i, j, k);
      const long il = LIDX(D.g, i, j, k);
      D.A.c[il] = G.A.c[ig];
      D.A.u[0][il] = G.A.u[0][ig];
      D.A.u[1][il] = G.A.u[1][ig];
      D.A.u[2][il] = G.A.u[2][ig];
      if (!last) { D.p1[il] = G.p1[ig]; D.p2[il] = G.p2[ig]; }
      if (top) D.b[il] = bin ? (T)bin[ig] : G.b[ig];
    }
  }
  __syncthreads();
  vcycle_block<T, false>(lv, 0, nl, pre, post, symmetric);
  const LevelDev<T> &G = lev[lc];
  const LevelDev<T> &D = lv[0];
  if (xout) {
    FOR_POINTS(G.g, xout[LIDX(G.g, i, j, k)] = (double)D.x[LIDX(D.g, i, j, k)]);
  } else {
    FOR_POINTS(G.g, G.x[LIDX(G.g, i, j, k)] = D.x[LIDX(D.g, i, j, k)]);
  }
}

template <typename T>
__global__ __launch_bounds__(CWG) void k_coarse_vcycle_lds(const LevelDev<T> *__restrict__ lev,
                                                           int lc, int L, int pre, int post,
                                                           int symmetric)
{
  vcycle_lds<T>(lev, lc, L, pre, post, symmetric, nullptr, nullptr);
}

// LDS bytes of the bottom levels lc .. L-1 in k_coarse_vcycle_lds
static long lds_bottom_bytes(const int (*n)[3], int lc, int L, size_t es, size_t rec)
{
  long b = lds_level_table_bytes(L - lc, (long)rec);
  for (int l = lc; l < L; l++) b += 9L * (long)es * n[l][0] * n[l][1] * n[l][2];
  return b;
}

// coarse operators of levels ls .. L-2 -> ls+1 .. L-1
template <typename T, bool FULL>
__global__ __launch_bounds__(CWG) void k_coarse_setup(const LevelDev<T> *__restrict__ lev,
                                                      int ls, int L)
{
  for (int l = ls; l <= L - 2; l++) {
    const LevelDev<T> &F = lev[l], &Cc = lev[l + 1];
    FOR_POINTS(Cc.g, (coarse_point<T, FULL>(F.g, Cc.g, F.A, Cc.A, F.dir, F.o1, F.o2, F.p1, F.p2,
                                            i, j, k)));
  }
}

// ---- host side ----------------------------------------------------------------
struct pfhip_mg {
  pfhip_problem *pb;
  int L, prec, pre, post, symmetric;
  int lc;                      // first level run by the single-workgroup kernels
  int n[MAXL][3], dir[MAXL];
  int full;                    // full storage (FullJacobian PC matrix)
  void *A[MAXL][7];            // C, E, N, U and (full) W, S, L
  void *x[MAXL], *b[MAXL], *t[MAXL], *p1[MAXL], *p2[MAXL];
  void *xa[MAXL];              // second iterate buffer of the grid levels (l < lc)
  void *raw[MAXL][13];         // allocations; each array starts one element in, so
                               // the ghosted index of an even x is vector aligned
  int fuse;                    // red -> black pairs as one launch (k_rbgs_pair)
  int rrz;                     // fused level residual + z restriction
  int cz;                      // z-marching coarse operators (k_coarse_z)
  int pv[MAXL];                // pair tile shape of each grid level (tune_pairs)
  int pv_din, pv_dout;         // level 0's shapes for its FP64-input and FP64-result pairs
  int a16;                     // bit 0 / 1: the V-cycle's FP64 input / result vector has
                               // its even-i points 16-byte aligned (set per call)
  int gx_a16;                  // a16 of the vectors the graph gx was captured for
  double bscale;               // the V-cycle's right-hand side is bscale * b (set per call)
  const double *dq_u;          // non-null: also the J.v sums u.x, x.x, |x|_1 of the result
                               // (into the problem's DQ slots), u = dq_u (set per call)
  int dq_done;                 // the sequence's last pair gathered them
  int gx_dq;                   // the graph gx gathers them
  double *dqp;                 // their per-workgroup partials
  long dqp_n;                  // partials per sum that dqp holds
  int tuned;                   // pv chosen (at the first V-cycle)
  void *levdev;                // device LevelDev<T>[L]
  long lds_bytes;              // > 0: the bottom runs in LDS (k_coarse_vcycle_lds)
  int use_graph;               // replay the captured V-cycle (mg_capture)
  hipStream_t cs;              // capture stream
  hipGraphExec_t gx;           // the captured V-cycle
  void **io;                   // device: the replay's right-hand side and result
};

template <typename T> static int tune_pairs(pfhip_mg *mg, hipStream_t st);
static int rbgs_forced();

template <typename T> static Mat<T> matof(pfhip_mg *mg, int l)
{
  Mat<T> M;
  M.c = (T *)mg->A[l][0];
  for (int q = 0; q < 3; q++) {
    M.u[q] = (T *)mg->A[l][1 + q];
    M.l[q] = mg->full ? (T *)mg->A[l][4 + q] : nullptr;
  }
  return M;
}

static void offdirs(int d, int *o1, int *o2)
{
  // a3..a6 order of PL/mg_semi.c:470-492
  if (d == 0) { *o1 = 1; *o2 = 2; }
  else if (d == 1) { *o1 = 0; *o2 = 2; }
  else { *o1 = 1; *o2 = 0; }
}

template <typename T> static int upload_levels(pfhip_mg *mg)
{
  LevelDev<T> h[MAXL];
  memset(h, 0, sizeof(h));
  for (int l = 0; l < mg->L; l++) {
    h[l].g = lgeom(mg->n[l]);
    h[l].dir = (l < mg->L - 1) ? mg->dir[l] : 0;
    offdirs(h[l].dir, &h[l].o1, &h[l].o2);
    h[l].A = matof<T>(mg, l);
    h[l].x = (T *)mg->x[l]; h[l].b = (T *)mg->b[l]; h[l].t = (T *)mg->t[l];
    h[l].p1 = (T *)mg->p1[l]; h[l].p2 = (T *)mg->p2[l];
  }
  int rc = pfhip_malloc(&mg->levdev, (long)sizeof(LevelDev<T>) * mg->L);
  if (rc) return rc;
  HIPCHK(hipMemcpy(mg->levdev, h, sizeof(LevelDev<T>) * mg->L, hipMemcpyHostToDevice));
  return PFHIP_OK;
}

static int mg_create(pfhip_problem *pb, int max_levels, int pre, int post, int symmetric,
                     int precision, int full, pfhip_mg **out)
{
  if (precision != 64 && precision != 32) {
    pfhip_set_error("mg precision must be 64 or 32");
    return PFHIP_ERR_ARG;
  }
  if (pre < 0 || post < 0 || pre > MAX_SWEEPS || post > MAX_SWEEPS) {
    pfhip_set_error("sweep counts must lie in [0, %d]", MAX_SWEEPS);
    return PFHIP_ERR_ARG;
  }
  pfhip_mg *mg = (pfhip_mg *)calloc(1, sizeof(pfhip_mg));
  mg->pb = pb;
  mg->prec = precision;
  mg->pre = pre; mg->post = post; mg->symmetric = symmetric;
  mg->full = full;
  int N[3] = { pb->d.nx, pb->d.ny, pb->d.nz };
  double D[3] = { pb->d.dx, pb->d.dy, pb->d.dz };
  if (max_levels <= 0 || max_levels > MAXL) max_levels = MAXL;
  for (int q = 0; q < 3; q++) mg->n[0][q] = N[q];
  int l;
  for (l = 0; l < max_levels - 1; l++) {
    double min_spacing = D[0] + D[1] + D[2];
    int d = -1;
    for (int q = 0; q < 3; q++)
      if (N[q] >= 2 && D[q] < min_spacing) { min_spacing = D[q]; d = q; }
    if (d == -1) break;
    D[d] *= 2;
    N[d] = (N[d] + 1) / 2;
    mg->dir[l] = d;
    for (int q = 0; q < 3; q++) mg->n[l + 1][q] = N[q];
  }
  mg->L = l + 1;
  if (mg->L < 2) {
    free(mg);
    pfhip_set_error("grid too small for a V-cycle");
    return PFHIP_ERR_ARG;
  }
  // first level small enough for the single-workgroup bottom (never level 0)
  const char *fe = getenv("PFHIP_MG_FUSE");
  mg->fuse = (fe && *fe) ? atoi(fe) : 1;
  const char *fr = getenv("PFHIP_MG_RRZ");
  mg->rrz = (fr && *fr) ? atoi(fr) : 1;
  const char *fc = getenv("PFHIP_MG_CZ");
  mg->cz = (fc && *fc) ? atoi(fc) : 1;
  long small = MG_SMALL_DEFAULT;
  const char *env = getenv("PFHIP_MG_SMALL");
  if (env && *env) small = atol(env);
  mg->lc = mg->L - 1;
  for (l = 1; l < mg->L; l++)
    if (npts(mg->n[l]) <= small) { mg->lc = l; break; }
  size_t es = (precision == 64) ? sizeof(double) : sizeof(float);
  // bottom in LDS (symmetric storage): the first level from which every
  // coarser level fits the LDS budget, if that is no larger than the
  // single-workgroup bound
  const char *fl = getenv("PFHIP_MG_LDS");
  const long lds_budget = (fl && *fl) ? atol(fl) : MG_LDS_DEFAULT;
  mg->lds_bytes = 0;
  int dev = 0, lds_max = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess)
    lds_max = 0;
  if (!full && lds_budget > 0 && lds_budget <= lds_max && !(env && *env)) {
    const size_t rec = (precision == 64) ? sizeof(LevelDev<double>) : sizeof(LevelDev<float>);
    for (l = 1; l <= mg->L - 2; l++) {
      const long bytes = lds_bottom_bytes(mg->n, l, mg->L, es, rec);
      if (bytes <= lds_budget && npts(mg->n[l]) <= small) {
        mg->lc = l;
        mg->lds_bytes = bytes;
        break;
      }
    }
    // more than the default dynamic LDS of a launch: raise the kernel's limit
    if (mg->lds_bytes > 0) {
      hipError_t e = (precision == 64)
          ? hipFuncSetAttribute((const void *)k_coarse_vcycle_lds<double>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)mg->lds_bytes)
          : hipFuncSetAttribute((const void *)k_coarse_vcycle_lds<float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)mg->lds_bytes);
      if (e != hipSuccess) {
        // keep the bottom in global memory with the single-workgroup bound
        (void)hipGetLastError();
        mg->lds_bytes = 0;
        mg->lc = mg->L - 1;
        for (l = 1; l < mg->L; l++)
          if (npts(mg->n[l]) <= small) { mg->lc = l; break; }
      }
    }
  }
  int rc = PFHIP_OK;
  for (l = 0; l < mg->L && rc == PFHIP_OK; l++) {
    long s = lsize(mg->n[l]);
    void **ptrs[13] = { &mg->A[l][0], &mg->A[l][1], &mg->A[l][2], &mg->A[l][3],
                        &mg->x[l], &mg->b[l], &mg->t[l], &mg->p1[l], &mg->p2[l],
                        &mg->xa[l], &mg->A[l][4], &mg->A[l][5], &mg->A[l][6] };
    for (int q = 0; q < 13 && rc == PFHIP_OK; q++) {
      if ((q == 9 && l >= mg->lc) || (q >= 10 && !full)) continue;
      rc = pfhip_malloc(&mg->raw[l][q], (long)((s + 1) * es));
      if (rc == PFHIP_OK) {
        *ptrs[q] = (char *)mg->raw[l][q] + es;
        hipError_t e = hipMemset(mg->raw[l][q], 0, (s + 1) * es);
        if (e != hipSuccess) { pfhip_set_error("hipMemset"); rc = PFHIP_ERR_HIP; }
      }
    }
  }
  if (rc == PFHIP_OK)
    rc = (precision == 64) ? upload_levels<double>(mg) : upload_levels<float>(mg);
  if (rc != PFHIP_OK) { pfhip_mg_destroy(mg); return rc; }
  int forced = rbgs_forced();
  for (l = 0; l < MAXL; l++) mg->pv[l] = (forced >= 0) ? forced : 0;
  mg->pv_din = mg->pv_dout = (forced >= 0) ? forced : 0;
  mg->tuned = (forced >= 0) || !mg->fuse || full;
  const char *fg = getenv("PFHIP_MG_GRAPH");
  mg->use_graph = (fg && *fg) ? atoi(fg) : 1;
  HIPCHK(hipDeviceSynchronize());
  *out = mg;
  return PFHIP_OK;
}

extern "C" int pfhip_mg_create(pfhip_problem *pb, int max_levels, int pre, int post,
                               int symmetric, int precision, pfhip_mg **out)
{
  return mg_create(pb, max_levels, pre, post, symmetric, precision, 0, out);
}

extern "C" int pfhip_mg_create_full(pfhip_problem *pb, int max_levels, int pre, int post,
                                    int symmetric, int precision, pfhip_mg **out)
{
  return mg_create(pb, max_levels, pre, post, symmetric, precision, 1, out);
}

extern "C" int pfhip_mg_destroy(pfhip_mg *mg)
{
  if (!mg) return PFHIP_OK;
  if (mg->gx) (void)hipGraphExecDestroy(mg->gx);
  if (mg->cs) (void)hipStreamDestroy(mg->cs);
  if (mg->io) pfhip_free(mg->io);
  if (mg->dqp) pfhip_free(mg->dqp);
  for (int l = 0; l < mg->L; l++) {
    for (int q = 0; q < 13; q++) if (mg->raw[l][q]) hipFree(mg->raw[l][q]);
  }
  if (mg->levdev) hipFree(mg->levdev);
  free(mg);
  return PFHIP_OK;
}

extern "C" int pfhip_mg_num_levels(pfhip_mg *mg, int *levels)
{
  *levels = mg->L;
  return PFHIP_OK;
}

template <typename T> static int mg_coarse_setup_t(pfhip_mg *mg, hipStream_t st);

template <typename T>
static int mg_setup_t(pfhip_mg *mg, const double *const *Asym, hipStream_t st)
{
  LGeom g0 = lgeom(mg->n[0]);
  const double *src[7] = { Asym[0], Asym[2], Asym[4], Asym[6], Asym[1], Asym[3], Asym[5] };
  for (int q = 0; q < (mg->full ? 7 : 4); q++) {
    hipLaunchKernelGGL((k_convert<double, T>), dim3(ntiles(g0)), LBLOCK, 0, st, g0, src[q],
                       (T *)mg->A[0][q]);
    LAUNCHCHK();
  }
  return mg_coarse_setup_t<T>(mg, st);
}

// coarse operators of every level from the finest-level matrix
template <typename T> static int mg_coarse_setup_t(pfhip_mg *mg, hipStream_t st)
{
  // large levels: grid kernels down to (and into) level lc
  for (int l = 0; l < mg->lc; l++) {
    LGeom gf = lgeom(mg->n[l]), gc = lgeom(mg->n[l + 1]);
    int d = mg->dir[l], o1, o2;
    offdirs(d, &o1, &o2);
    if (!mg->full && mg->cz && d == 2) {
      const unsigned nb = (unsigned)((gc.n[0] + 63) / 64) * ((gc.n[1] + 3) / 4);
      hipLaunchKernelGGL(k_coarse_z<T>, dim3((nb + 7u) & ~7u), LBLOCK, 0, st, gf, gc,
                         matof<T>(mg, l), matof<T>(mg, l + 1), (T *)mg->p1[l], (T *)mg->p2[l]);
    } else if (mg->full)
      hipLaunchKernelGGL((k_coarse_fused<T, true>), dim3(ntiles(gc)), LBLOCK, 0, st, gf, gc,
                         matof<T>(mg, l), matof<T>(mg, l + 1), d, o1, o2, (T *)mg->p1[l],
                         (T *)mg->p2[l]);
    else
      hipLaunchKernelGGL((k_coarse_fused<T, false>), dim3(ntiles(gc)), LBLOCK, 0, st, gf, gc,
                         matof<T>(mg, l), matof<T>(mg, l + 1), d, o1, o2, (T *)mg->p1[l],
                         (T *)mg->p2[l]);
    LAUNCHCHK();
  }
  // small levels: one workgroup
  if (mg->lc <= mg->L - 2) {
    if (mg->full)
      hipLaunchKernelGGL((k_coarse_setup<T, true>), dim3(1), dim3(CWG), 0, st,
                         (const LevelDev<T> *)mg->levdev, mg->lc, mg->L);
    else
      hipLaunchKernelGGL((k_coarse_setup<T, false>), dim3(1), dim3(CWG), 0, st,
                         (const LevelDev<T> *)mg->levdev, mg->lc, mg->L);
    LAUNCHCHK();
  }
  return PFHIP_OK;
}

// PC setup of KinsolPC (PL/kinsol_pc.c:147-154) without an intermediate
// matrix: the Jacobian of pressure p (its symmetric part, or for a full-
// storage hierarchy the full Jacobian with the surface couplings) is
// assembled straight into the finest level (in the level precision), then
// the coarse operators.
int pfhip_mg_setup_jacobian(pfhip_mg *mg, const double *p, double dt, void *stream)
{
  void *J[7] = { mg->A[0][0], mg->A[0][4], mg->A[0][1], mg->A[0][5], mg->A[0][2],
                 mg->A[0][6], mg->A[0][3] };
  int rc = pfhip_jacobian_into(mg->pb, p, dt, !mg->full, J, mg->prec == 32, stream);
  if (rc) return rc;
  return (mg->prec == 64) ? mg_coarse_setup_t<double>(mg, STREAM(stream))
                          : mg_coarse_setup_t<float>(mg, STREAM(stream));
}

extern "C" int pfhip_mg_setup(pfhip_mg *mg, const double *const *Asym, void *stream)
{
  if (!Asym || !Asym[0] || !Asym[2] || !Asym[4] || !Asym[6]) {
    pfhip_set_error("mg_setup: need C, E, N, U arrays");
    return PFHIP_ERR_ARG;
  }
  if (mg->full && (!Asym[1] || !Asym[3] || !Asym[5])) {
    pfhip_set_error("mg_setup: a full-storage multigrid needs all seven arrays");
    return PFHIP_ERR_ARG;
  }
  return (mg->prec == 64) ? mg_setup_t<double>(mg, Asym, STREAM(stream))
                          : mg_setup_t<float>(mg, Asym, STREAM(stream));
}

// RedBlackGSPoint call sequence (PL/rb_GS_point.c:100-300): the list of
// half-sweeps (parity, initial) it performs, and whether it first zeroes x
struct HalfSweep { int par, init; };
static int sweep_list(int symmetric, int zero, int max_iter, HalfSweep *seq, int *fill)
{
  int n = 0, count = symmetric ? 1 : 2, rb = 0;   // red first; red = odd parity
  *fill = 0;
  if (zero) {
    count++;
    if ((count / 2) > max_iter) {
      *fill = 1;
      if (!symmetric) return 0;
    }
    seq[n].par = 1; seq[n].init = 1; n++;
    rb = !rb;
  }
  for (; (count / 2) <= max_iter; count++) {
    seq[n].par = rb ? 0 : 1; seq[n].init = 0; n++;
    rb = !rb;
  }
  return n;
}

// Fused pair tile shapes (rows TY of a tile, z-chunk KC).  Each multigrid
// times them per grid level when it is created (tune_pairs) and keeps the
// fastest; PFHIP_RBGS_VARIANT=v forces shape v on every level.
#define NPAIRV 17
// the shapes tuned in each precision: the deeper-prefetch forms (v >= 10)
// are FP32 only (an FP64 plane set is twice the registers)
template <typename T> static int npairv() { return sizeof(T) == 4 ? NPAIRV : 10; }
static int rbgs_forced()
{
  const char *e = getenv("PFHIP_RBGS_VARIANT");
  int v = (e && *e) ? atoi(e) : -1;
  return (v >= 0 && v < NPAIRV) ? v : -1;
}

template <typename T, int TY, int KC>
static void launch_pair_t(const LGeom &g, const Mat<T> &A, const T *xin, T *xout, const T *b,
                          int init, hipStream_t st)
{
  unsigned gx = (g.n[0] + PX - 1) / PX, gy = (g.n[1] + TY - 1) / TY, gz = (g.n[2] + KC - 1) / KC;
  unsigned nb = gx * gy * gz, nb8 = (nb + 7u) & ~7u;
  if (init)
    hipLaunchKernelGGL((k_rbgs_pair<T, TY, KC, true>), dim3(nb8), dim3(64, TY + 2, 1), 0, st, g,
                       A, xin, xout, b);
  else
    hipLaunchKernelGGL((k_rbgs_pair<T, TY, KC, false>), dim3(nb8), dim3(64, TY + 2, 1), 0, st, g,
                       A, xin, xout, b);
}

// returns the launch's workgroup count (the DQ partials it wrote)
template <typename T, int TY, int KC, bool YL, int PF = 1>
static int launch_pair2_t(const LGeom &g, const Mat<T> &A, const T *xin, T *xout, const T *b,
                          int init, hipStream_t st, double *dout, const double *din,
                          void *const *io, bool a16, double dsc, const double *dqu,
                          double *dqp)
{
  unsigned gx = (g.n[0] + PX2 - 1) / PX2, gy = (g.n[1] + TY - 1) / TY, gz = (g.n[2] + KC - 1) / KC;
  unsigned nb = gx * gy * gz, nb8 = (nb + 7u) & ~7u;
  const dim3 grid(nb8), blk(64, TY + 2, 1);
#define P2(INIT, DOUT, DIN, A16, DQ, DO, DI, BO, DS)                                          \
  hipLaunchKernelGGL((k_rbgs_pair2<T, TY, KC, INIT, DOUT, DIN, YL, PF, A16, DQ>), grid, blk, 0, \
                     st, g, A, xin, xout, b, DO, DI, BO, io, DS, dqu, dqp)
  if (init && din && a16) P2(true, false, true, true, false, nullptr, din, (T *)b, dsc);
  else if (init && din) P2(true, false, true, false, false, nullptr, din, (T *)b, dsc);
  else if (init) P2(true, false, false, false, false, nullptr, nullptr, nullptr, 1.0);
  else if (dout && dqp && a16) P2(false, true, false, true, true, dout, nullptr, nullptr, 1.0);
  else if (dout && dqp) P2(false, true, false, false, true, dout, nullptr, nullptr, 1.0);
  else if (dout && a16) P2(false, true, false, true, false, dout, nullptr, nullptr, 1.0);
  else if (dout) P2(false, true, false, false, false, dout, nullptr, nullptr, 1.0);
  else P2(false, false, false, false, false, nullptr, nullptr, nullptr, 1.0);
#undef P2
  return (int)nb8;
}

// shape v % 5: (TY, KC) = (14, 16), (6, 16), (6, 32), (14, 8), (14, 4);
// v >= 5 takes the y neighbours through LDS (the YL form; even x extents);
// v = 10..14: the YL form loading two planes ahead (FP32), 15 and 16:
// three planes ahead with (14, 16) and (14, 8);
// dout (even x extents, non-initial pairs only) takes the result in FP64,
// din (even x extents, initial pairs only) gives the right-hand side in
// FP64 and b receives it in the level precision; a16: that FP64 vector's
// even-i points are 16-byte aligned; dqp (with dout): the result's J.v
// partials (k_rbgs_pair2's DQ).  Returns the workgroup count of a
// two-points-per-lane launch, 0 for the one-point form.
template <typename T>
static int launch_pair(const LGeom &g, const Mat<T> &A, const T *xin, T *xout, const T *b,
                        int init, int v, hipStream_t st, double *dout = nullptr,
                        const double *din = nullptr, void *const *io = nullptr,
                        bool a16 = false, double dsc = 1.0, const double *dqu = nullptr,
                        double *dqp = nullptr)
{
  if ((g.n[0] % 2) == 0) {
    // two points per lane (aligned vector loads need an even x extent)
    switch (v) {
      case 1: return launch_pair2_t<T, 6, 16, false>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 2: return launch_pair2_t<T, 6, 32, false>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 3: return launch_pair2_t<T, 14, 8, false>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 4: return launch_pair2_t<T, 14, 4, false>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 5: return launch_pair2_t<T, 14, 16, true>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 6: return launch_pair2_t<T, 6, 16, true>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 7: return launch_pair2_t<T, 6, 32, true>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 8: return launch_pair2_t<T, 14, 8, true>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 9: return launch_pair2_t<T, 14, 4, true>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      default: break;
    }
    if constexpr (sizeof(T) == 4) switch (v) {
      case 10: return launch_pair2_t<T, 14, 16, true, 2>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 11: return launch_pair2_t<T, 6, 16, true, 2>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 12: return launch_pair2_t<T, 6, 32, true, 2>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 13: return launch_pair2_t<T, 14, 8, true, 2>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 14: return launch_pair2_t<T, 14, 4, true, 2>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 15: return launch_pair2_t<T, 14, 16, true, 3>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      case 16: return launch_pair2_t<T, 14, 8, true, 3>(g, A, xin, xout, b, init, st, dout, din, io, a16, dsc, dqu, dqp);
      default: break;
    }
    return launch_pair2_t<T, 14, 16, false>(g, A, xin, xout, b, init, st, dout, din, io, a16,
                                            dsc, dqu, dqp);
  }
  switch (v % 5) {
    case 1: launch_pair_t<T, 6, 16>(g, A, xin, xout, b, init, st); break;
    case 2: launch_pair_t<T, 6, 32>(g, A, xin, xout, b, init, st); break;
    case 3: launch_pair_t<T, 14, 8>(g, A, xin, xout, b, init, st); break;
    case 4: launch_pair_t<T, 14, 4>(g, A, xin, xout, b, init, st); break;
    default: launch_pair_t<T, 14, 16>(g, A, xin, xout, b, init, st); break;
  }
  return 0;
}

// The autotuner, run at the first V-cycle of a hierarchy (so on its real
// operator, and never for hierarchies that the ensemble drives a workgroup
// at a time): every pair shape on every grid level, one untimed launch and
// then TUNE_REPS timed ones (HIP events), keeping the fastest per level.
// Level 0 also times the two forms that touch the solver's FP64 vectors
// (the first pair reading the right-hand side bin, the last writing the
// result into xout: their extra bytes shift the best shape), with the
// vectors of this first V-cycle, which then overwrites what they wrote.
// The iterate buffers are re-zeroed afterwards.
#define TUNE_REPS 3
template <typename T>
static int tune_pairs(pfhip_mg *mg, hipStream_t st, const double *bin, double *xout)
{
  mg->tuned = 1;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int l = 0; l < mg->lc; l++) {
    if (!mg->xa[l]) continue;
    LGeom g = lgeom(mg->n[l]);
    Mat<T> A = matof<T>(mg, l);
    const bool fp64_forms = l == 0 && bin && xout && (g.n[0] % 2) == 0;
    for (int form = 0; form < (fp64_forms ? 3 : 1); form++) {
      // form 0: level-precision operands; 1: FP64 right-hand side (zero
      // initial guess); 2: FP64 result
      const int init = form == 1;
      double *dout = form == 2 ? xout : nullptr;
      const double *din = form == 1 ? bin : nullptr;
      int *pick = form == 0 ? &mg->pv[l] : form == 1 ? &mg->pv_din : &mg->pv_dout;
      const bool a16 = form == 1 ? (mg->a16 & 1) : form == 2 ? (mg->a16 & 2) : false;
      // the result form as the V-cycles will run it: with the J.v sums
      // when this hierarchy gathers them
      const double *dqu = form == 2 ? mg->dq_u : nullptr;
      double *dqp = dqu ? mg->dqp : nullptr;
      float best = 0.f;
      for (int v = 0; v < npairv<T>(); v++) {
        launch_pair<T>(g, A, (const T *)mg->x[l], (T *)mg->xa[l], (const T *)mg->b[l], init, v,
                       st, dout, din, nullptr, a16, mg->bscale, dqu, dqp);
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < TUNE_REPS; r++)
          launch_pair<T>(g, A, (const T *)mg->x[l], (T *)mg->xa[l], (const T *)mg->b[l], init,
                         v, st, dout, din, nullptr, a16, mg->bscale, dqu, dqp);
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (v == 0 || ms < best) { best = ms; *pick = v; }
      }
    }
    const long sz = lsize(mg->n[l]) + 1;
    hipLaunchKernelGGL(k_fill_t<T>, dim3(256), dim3(256), 0, st, (T *)mg->raw[l][4], sz);
    hipLaunchKernelGGL(k_fill_t<T>, dim3(256), dim3(256), 0, st, (T *)mg->raw[l][9], sz);
    LAUNCHCHK();
  }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  return PFHIP_OK;
}

extern "C" int pfhip_mg_pair_shapes(pfhip_mg *mg, int *shape, int max)
{
  int n = mg->lc < max ? mg->lc : max;
  for (int l = 0; l < n; l++) shape[l] = mg->pv[l];
  return n;
}

// level 0's shapes for the pair reading the FP64 right-hand side and the
// pair writing the FP64 result
extern "C" int pfhip_mg_pair_shapes_fp64(pfhip_mg *mg, int *din, int *dout)
{
  *din = mg->pv_din;
  *dout = mg->pv_dout;
  return PFHIP_OK;
}

// smoothing of grid level l: red -> black pairs run fused (x and its second
// buffer swap roles after each pair), any other half-sweep runs alone.
// With dout, a last half-sweep pair on an even x extent writes its result
// into dout in FP64 (and *wrote is set); x then holds no valid iterate.
template <typename T>
static int rbgs(pfhip_mg *mg, int l, int zero, int max_iter, hipStream_t st,
                double *dout = nullptr, int *wrote = nullptr, const double *din = nullptr,
                void *const *io = nullptr)
{
  LGeom g = lgeom(mg->n[l]);
  dim3 grid(ntiles(g));
  Mat<T> A = matof<T>(mg, l);
  const T *b = (const T *)mg->b[l];
  HalfSweep seq[2 * MAX_SWEEPS + 4];
  int fill;
  int n = sweep_list(mg->symmetric, zero, max_iter, seq, &fill);
  if (fill) {
    hipLaunchKernelGGL(k_fill_t<T>, dim3(256), dim3(256), 0, st, (T *)mg->x[l], lsize(mg->n[l]));
    LAUNCHCHK();
  }
  for (int q = 0; q < n;) {
    int dq_nb = 0;
    const bool pair = mg->fuse && !mg->full && mg->xa[l] && q + 1 < n && seq[q].par == 1 &&
                      seq[q + 1].par == 0;
    double *d = (pair && dout && q + 2 == n && !seq[q].init && (g.n[0] % 2) == 0) ? dout
                                                                                 : nullptr;
    const double *di = (pair && din && q == 0 && seq[q].init) ? din : nullptr;
    // level 0's launches, timed by form: level-precision operands, the pair
    // writing the FP64 result (4 more bytes stored per point) and the pair
    // converting the FP64 right-hand side (4 more read, 4 more stored)
    const int mark = (l != 0) ? -1 : d ? PFHIP_PROF_RBGS_DOUT : di ? PFHIP_PROF_RBGS_DIN
                                                                   : PFHIP_PROF_RBGS;
    if (mark >= 0) pfhip_prof_mark(mark, 0, st);
    if (pair) {
      const int v = (l == 0 && d) ? mg->pv_dout : (l == 0 && di) ? mg->pv_din : mg->pv[l];
      const bool a16 = d ? (mg->a16 & 2) : di ? (mg->a16 & 1) : false;
      // the result's J.v sums from the pair that writes it
      const bool dq = d && mg->dq_u;
      dq_nb = launch_pair<T>(g, A, (const T *)mg->x[l], (T *)mg->xa[l], b, seq[q].init, v, st,
                             d, di, io, a16, di ? mg->bscale : 1.0, dq ? mg->dq_u : nullptr,
                             dq ? mg->dqp : nullptr);
      if (dq) mg->dq_done = 1;
      if (d && wrote) *wrote = 1;
      void *tmp = mg->x[l]; mg->x[l] = mg->xa[l]; mg->xa[l] = tmp;
      q += 2;
    } else {
      if (mg->full)
        hipLaunchKernelGGL((k_rbgs<T, true>), grid, LBLOCK, 0, st, g, A, (T *)mg->x[l], b,
                           seq[q].par, seq[q].init);
      else
        hipLaunchKernelGGL((k_rbgs<T, false>), grid, LBLOCK, 0, st, g, A, (T *)mg->x[l], b,
                           seq[q].par, seq[q].init);
      q += 1;
    }
    if (mark >= 0) pfhip_prof_mark(mark, 1, st);
    LAUNCHCHK();
    if (dq_nb > 0 && mg->dq_done && d) {
      hipLaunchKernelGGL(k_dq_final, dim3(1), dim3(256), 0, st, mg->dqp, dq_nb,
                         mg->pb->red_dev + PFHIP_SLOT_DQ);
      LAUNCHCHK();
    }
  }
  return PFHIP_OK;
}

template <typename T>
static void launch_resid_restrict_z(const LGeom &gf, const LGeom &gc, const Mat<T> &A,
                                    const T *x, const T *b, T *rc, const T *p1, const T *p2,
                                    hipStream_t st)
{
  constexpr int TY = 4, KC = 16;
  const int gx = (gf.n[0] + PX2 - 1) / PX2, gy = (gf.n[1] + TY - 1) / TY;
  const int gz = (gf.n[2] + KC - 1) / KC;
  const unsigned nb8 = ((unsigned)(gx * gy * gz) + 7u) & ~7u;
  hipLaunchKernelGGL((k_resid_restrict_z<T, TY, KC>), dim3(nb8), dim3(64, TY, 1), 0, st, gf,
                     gc, A, x, b, rc, p1, p2);
}

// V-cycle of PL/mg_semi.c:205-330 (MaxIter 1, zero initial guess)
// The launch sequence of one V-cycle.  io != nullptr: the FP64 interface
// vectors are not bin / xout but the addresses in io[0] / io[1] (device
// memory), read by the kernels that touch them, so the sequence can be
// captured once and replayed for other vectors.
template <typename T>
static int mg_apply_seq(pfhip_mg *mg, const double *bin, double *xout, hipStream_t st,
                        void *const *io)
{
  int L = mg->L, lc = mg->lc, rc, wrote = 0;
  LGeom g0 = lgeom(mg->n[0]);
  mg->dq_done = 0;
  // the first pre-smoothing pair of level 0 converts b itself when it is an
  // initial two-points-per-lane pair
  HalfSweep seq0[2 * MAX_SWEEPS + 4];
  int fill0;
  const int n0 = sweep_list(mg->symmetric, 1, mg->pre, seq0, &fill0);
  const bool din = mg->fuse && !mg->full && mg->xa[0] && (g0.n[0] % 2) == 0 && !fill0 &&
                   n0 >= 2 && seq0[0].par == 1 && seq0[0].init && seq0[1].par == 0;
  if (!din) {
    if (io)
      hipLaunchKernelGGL((k_convert_io<double, T>), dim3(ntiles(g0)), LBLOCK, 0, st, g0, bin,
                         (T *)mg->b[0], io, 0);
    else
      hipLaunchKernelGGL(k_convert_in<T>, dim3(ntiles(g0)), LBLOCK, 0, st, g0, bin,
                         (T *)mg->b[0], mg->bscale);
    LAUNCHCHK();
  }
  int l;
  for (l = 0; l < lc; l++) {
    if ((rc = rbgs<T>(mg, l, 1, mg->pre, st, nullptr, nullptr, (l == 0 && din) ? bin : nullptr,
                      io)))
      return rc;
    LGeom gf = lgeom(mg->n[l]), gc = lgeom(mg->n[l + 1]);
    if (mg->rrz && !mg->full && mg->dir[l] == 2 && (gf.n[0] & 1) == 0) {
      launch_resid_restrict_z<T>(gf, gc, matof<T>(mg, l), (const T *)mg->x[l],
                                 (const T *)mg->b[l], (T *)mg->b[l + 1], (const T *)mg->p1[l],
                                 (const T *)mg->p2[l], st);
      LAUNCHCHK();
      continue;
    }
    if (mg->full)
      hipLaunchKernelGGL((k_resid_restrict<T, true>), dim3(ntiles(gc)), LBLOCK, 0, st, gf, gc,
                         mg->dir[l], matof<T>(mg, l), (const T *)mg->x[l], (const T *)mg->b[l],
                         (T *)mg->b[l + 1], (const T *)mg->p1[l], (const T *)mg->p2[l]);
    else
      hipLaunchKernelGGL((k_resid_restrict<T, false>), dim3(ntiles(gc)), LBLOCK, 0, st, gf, gc,
                         mg->dir[l], matof<T>(mg, l), (const T *)mg->x[l], (const T *)mg->b[l],
                         (T *)mg->b[l + 1], (const T *)mg->p1[l], (const T *)mg->p2[l]);
    LAUNCHCHK();
  }
  // bottom of the cycle: levels lc .. L-1 in one workgroup; the coarsest
  // level's solve is RedBlackGSPoint with MaxIter 1
  if (lc <= L - 2 && mg->lds_bytes > 0) {
    hipLaunchKernelGGL(k_coarse_vcycle_lds<T>, dim3(1), dim3(CWG), (unsigned)mg->lds_bytes, st,
                       (const LevelDev<T> *)mg->levdev, lc, L, mg->pre, mg->post,
                       mg->symmetric);
    LAUNCHCHK();
  } else if (lc <= L - 2) {
    if (mg->full)
      hipLaunchKernelGGL((k_coarse_vcycle<T, true>), dim3(1), dim3(CWG), 0, st,
                         (const LevelDev<T> *)mg->levdev, lc, L, mg->pre, mg->post,
                         mg->symmetric);
    else
      hipLaunchKernelGGL((k_coarse_vcycle<T, false>), dim3(1), dim3(CWG), 0, st,
                         (const LevelDev<T> *)mg->levdev, lc, L, mg->pre, mg->post,
                         mg->symmetric);
    LAUNCHCHK();
  } else {
    if ((rc = rbgs<T>(mg, L - 1, 1, 1, st))) return rc;
  }
  for (l = lc - 1; l >= 0; l--) {
    LGeom gf = lgeom(mg->n[l]), gc = lgeom(mg->n[l + 1]);
    if (mg->cz && mg->dir[l] == 2) {
      const unsigned nb = (unsigned)((gf.n[0] + 63) / 64) * ((gf.n[1] + 3) / 4);
      hipLaunchKernelGGL(k_prolong_z<T>, dim3((nb + 7u) & ~7u), LBLOCK, 0, st, gf, gc,
                         (const T *)mg->x[l + 1], (T *)mg->x[l], (const T *)mg->p1[l],
                         (const T *)mg->p2[l]);
    } else {
      hipLaunchKernelGGL(k_prolong_add<T>, dim3(ntiles(gf)), LBLOCK, 0, st, gf, gc, mg->dir[l],
                         (const T *)mg->x[l + 1], (T *)mg->x[l], (const T *)mg->p1[l],
                         (const T *)mg->p2[l]);
    }
    LAUNCHCHK();
    if ((rc = rbgs<T>(mg, l, 0, mg->post, st, l == 0 ? xout : nullptr, &wrote, nullptr, io)))
      return rc;
  }
  if (!wrote) {
    if (io)
      hipLaunchKernelGGL((k_convert_io<T, double>), dim3(ntiles(g0)), LBLOCK, 0, st, g0,
                         (const T *)mg->x[0], xout, io, 1);
    else
      hipLaunchKernelGGL((k_convert<T, double>), dim3(ntiles(g0)), LBLOCK, 0, st, g0,
                         (const T *)mg->x[0], xout);
    LAUNCHCHK();
  }
  if (mg->dq_u && !mg->dq_done) {
    // the J.v sums of the result in a pass of their own
    const long rows = (long)g0.n[1] * g0.n[2];
    const int nb = (int)(rows < DQ_ROW_BLOCKS ? rows : DQ_ROW_BLOCKS);
    hipLaunchKernelGGL(k_dq_rows, dim3(nb), dim3(256), 0, st, g0, mg->dq_u, xout, mg->dqp, io);
    hipLaunchKernelGGL(k_dq_final, dim3(1), dim3(256), 0, st, mg->dqp, nb,
                       mg->pb->red_dev + PFHIP_SLOT_DQ);
    LAUNCHCHK();
  }
  return PFHIP_OK;
}

// Capture the V-cycle's launch sequence once (after the autotuner has
// chosen its tile shapes) and instantiate it.  Nothing runs during the
// capture, but the host-side bookkeeping does: the grid levels' iterate
// buffers swap after every fused pair, so the pointer state is saved and
// restored, and a sequence that does not return every level's buffers to
// where they started (an odd number of pairs on a level) is not replayable:
// the hierarchy then keeps launching directly.
template <typename T>
static int mg_capture(pfhip_mg *mg)
{
  void *x0[MAXL], *xa0[MAXL];
  for (int l = 0; l < mg->L; l++) { x0[l] = mg->x[l]; xa0[l] = mg->xa[l]; }
  if (!mg->cs) HIPCHK(hipStreamCreateWithFlags(&mg->cs, hipStreamNonBlocking));
  if (!mg->io) {
    // the two vector addresses, the right-hand side's factor and u
    int rc = pfhip_malloc((void **)&mg->io, 4 * (long)sizeof(void *));
    if (rc) return rc;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(mg->cs, hipStreamCaptureModeRelaxed));
  // any non-null FP64 vectors select the interface forms; io supplies the
  // real addresses at each replay
  int rc = mg_apply_seq<T>(mg, (const double *)mg->io, (double *)mg->io, mg->cs, mg->io);
  hipError_t e = hipStreamEndCapture(mg->cs, &g);
  bool same = true;
  for (int l = 0; l < mg->L; l++) {
    same = same && mg->x[l] == x0[l] && mg->xa[l] == xa0[l];
    mg->x[l] = x0[l];
    mg->xa[l] = xa0[l];
  }
  if (rc || e != hipSuccess || !g || !same) {
    if (g) (void)hipGraphDestroy(g);
    mg->use_graph = 0;
    if (rc) return rc;
    if (e != hipSuccess) {
      pfhip_set_error("V-cycle graph capture: %s", hipGetErrorString(e));
      return PFHIP_ERR_HIP;
    }
    return PFHIP_OK;                 // not replayable: launch directly
  }
  e = hipGraphInstantiate(&mg->gx, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  mg->gx_a16 = mg->a16;
  mg->gx_dq = mg->dq_u != nullptr;
  if (e != hipSuccess) {
    mg->gx = nullptr;
    pfhip_set_error("V-cycle graph instantiation: %s", hipGetErrorString(e));
    return PFHIP_ERR_HIP;
  }
  return PFHIP_OK;
}

// V-cycle of PL/mg_semi.c:205-330 (MaxIter 1, zero initial guess).  The
// first call tunes the smoother and launches directly; later calls replay
// the captured sequence (PFHIP_MG_GRAPH=0 launches directly always), except
// while the smoother's per-launch timers are on (they time each launch).
// J.v partials the level-0 pair launches or k_dq_rows can write: the most
// workgroups of any tile shape, and DQ_ROW_BLOCKS
static long dq_partials(const pfhip_mg *mg)
{
  const int *n = mg->n[0];
  const int shapes[5][2] = { { 14, 16 }, { 6, 16 }, { 6, 32 }, { 14, 8 }, { 14, 4 } };
  long most = DQ_ROW_BLOCKS;
  for (int v = 0; v < 5; v++) {
    long nb = (long)((n[0] + PX2 - 1) / PX2) * ((n[1] + shapes[v][0] - 1) / shapes[v][0]) *
              ((n[2] + shapes[v][1] - 1) / shapes[v][1]);
    nb = (nb + 7) & ~7L;
    if (nb > most) most = nb;
  }
  return most;
}

template <typename T>
static int mg_apply_t(pfhip_mg *mg, const double *bin, double bscale, double *xout,
                      const double *dq_u, hipStream_t st)
{
  int rc;
  mg->bscale = bscale;
  mg->dq_u = dq_u;
  if (dq_u && !mg->dqp) {
    mg->dqp_n = dq_partials(mg);
    if ((rc = pfhip_malloc((void **)&mg->dqp, 3 * mg->dqp_n * (long)sizeof(double)))) return rc;
  }
  // an even i has an odd ghosted index, so its FP64 value pair is 16-byte
  // aligned when the vector starts 8 bytes past a 16-byte boundary
  mg->a16 = ((((uintptr_t)bin) & 15) == 8 ? 1 : 0) | ((((uintptr_t)xout) & 15) == 8 ? 2 : 0);
  if (!mg->tuned) {
    if ((rc = tune_pairs<T>(mg, st, bin, xout))) return rc;
    return mg_apply_seq<T>(mg, bin, xout, st, nullptr);
  }
  const bool timed = pfhip_prof_enabled(PFHIP_PROF_RBGS) ||
                     pfhip_prof_enabled(PFHIP_PROF_RBGS_DOUT) ||
                     pfhip_prof_enabled(PFHIP_PROF_RBGS_DIN);
  if (mg->use_graph && !timed) {
    // the captured launches read the vectors with the access width of the
    // vectors they were captured for: others are launched directly
    if (mg->gx && (mg->gx_a16 != mg->a16 || mg->gx_dq != (dq_u != nullptr)))
      return mg_apply_seq<T>(mg, bin, xout, st, nullptr);
    if (!mg->gx && (rc = mg_capture<T>(mg))) return rc;
    if (mg->gx) {
      hipLaunchKernelGGL(k_set_io, dim3(1), dim3(1), 0, st, mg->io, bin, xout, bscale, dq_u);
      LAUNCHCHK();
      HIPCHK(hipGraphLaunch(mg->gx, st));
      return PFHIP_OK;
    }
  }
  return mg_apply_seq<T>(mg, bin, xout, st, nullptr);
}

extern "C" int pfhip_mg_apply(pfhip_mg *mg, const double *b, double *x, void *stream)
{
  return (mg->prec == 64) ? mg_apply_t<double>(mg, b, 1.0, x, nullptr, STREAM(stream))
                          : mg_apply_t<float>(mg, b, 1.0, x, nullptr, STREAM(stream));
}

// x = M^{-1} (bscale * b): the V-cycle of the vector a scaling pass of b by
// bscale would have stored, with the factor applied where the V-cycle reads
// b (a Krylov basis vector kept unnormalised beside its 1/norm)
extern "C" int pfhip_mg_apply_scaled(pfhip_mg *mg, const double *b, double bscale, double *x,
                                     void *stream)
{
  return (mg->prec == 64) ? mg_apply_t<double>(mg, b, bscale, x, nullptr, STREAM(stream))
                          : mg_apply_t<float>(mg, b, bscale, x, nullptr, STREAM(stream));
}

// the same, and the sums the difference-quotient J.v of x needs (u.x, x.x
// and sum |x| over the grid's cells, u the Newton iterate) into the
// problem's device slots that pfhip_residual_jv_ready reads; no host
// synchronisation.  The V-cycle's last pair gathers them as it writes x
// where it writes x in FP64, a pass of their own otherwise.
extern "C" int pfhip_mg_apply_dq(pfhip_mg *mg, const double *b, double bscale, double *x,
                                 const double *u, void *stream)
{
  if (!u) {
    pfhip_set_error("mg_apply_dq: no iterate");
    return PFHIP_ERR_ARG;
  }
  return (mg->prec == 64) ? mg_apply_t<double>(mg, b, bscale, x, u, STREAM(stream))
                          : mg_apply_t<float>(mg, b, bscale, x, u, STREAM(stream));
}

// ---- ensemble multigrid ---------------------------------------------------------
// The levels of a member with at most ens_small() points (the bottom of its
// hierarchy; the whole hierarchy for small members) run inside one
// workgroup per member, one launch for every member with a request.  The
// larger levels of the larger members run as grid-level phases: one launch
// per phase and level covers that level's points of every member that has
// it (a half-sweep, the level residual, the restriction, the prolongation,
// the coarse operators), so no single member's one-workgroup V-cycle sets
// the cost of every heavy round.  A launch's blocks are dealt to members in
// segments (EnsLSeg: the member and its first block), found by a binary
// search over the segment starts.  The per-point arithmetic is the single-
// problem multigrid's (rbgs_point, resid_point, restrict_point,
// prolong_point, coarse_point), in the same sweep order.
struct EnsLSeg { int m, blk0; };

template <typename T> struct EnsLv {
  const LevelDev<T> *lev;
  const double *bin;    // this round's V-cycle right-hand side (FP64)
  double *xout;         // and result
  int L, lc, pre, post, symmetric, act;
};

// point sets a grid-level launch enumerates
enum { PS_FINE = 0, PS_HALF = 1, PS_COARSE = 2, PS_N = 3 };
enum { EL_CONV_IN, EL_FILL, EL_HALF, EL_RESTRICT, EL_PROLONG, EL_CONV_OUT, EL_COARSE };
#define ENS_LBLK 256

struct pfhip_ens_mgplan {
  int n, prec, G;                     // members, level precision, grid levels
  int *lc;                            // per member: first one-workgroup level
  int pre, post, symmetric;           // the members' common sweep schedule
  int nseg[MAXL][PS_N], nblk[MAXL][PS_N];
  EnsLSeg *dseg[MAXL][PS_N];          // device segment tables
  void *dM;                           // device member table EnsLv<T>[n], refilled per round
  int *dlist;                         // 0 .. n-1 (the one-workgroup kernels' members)
  int use_graph;                      // replay captured launch sequences
  hipStream_t cs;                     // capture stream
  hipGraphExec_t gx[2];               // [0] coarse operators, [1] V-cycle
  long glaunch[2];                    // kernels in each captured sequence
  long lds_bytes;                     // > 0: the one-workgroup part runs in LDS
  long launches;                      // kernel launches made so far
};

static long ens_small()
{
  const char *e = getenv("PFHIP_ENS_MG_SMALL");
  return (e && *e) ? atol(e) : MG_SMALL_DEFAULT;
}

template <typename T>
__global__ __launch_bounds__(ENS_LBLK) void k_ens_lvl(const EnsLSeg *__restrict__ seg, int nseg,
                                                      const EnsLv<T> *__restrict__ M, int l,
                                                      int OP, int par, int init)
{
  int lo = 0, hi = nseg - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  const EnsLv<T> &mm = M[seg[lo].m];
  if (!mm.act) return;
  const LevelDev<T> &F = mm.lev[l];
  const int idx = (b - seg[lo].blk0) * ENS_LBLK + threadIdx.x;
  int i, j, k;
  if (OP == EL_HALF) {
    const int nx = F.g.n[0], ny = F.g.n[1], nz = F.g.n[2];
    const int ncol = (nx + 1) >> 1;
    if (idx >= ncol * ny * nz) return;
    const int t = idx % ncol, r = idx / ncol;
    j = r % ny; k = r / ny;
    i = ((par - j - k) & 1) + 2 * t;
    if (i >= nx) return;
    rbgs_point<T, false>(F.g, F.A, F.x, F.b, par, init, i, j, k);
    return;
  }
  if (OP == EL_RESTRICT || OP == EL_COARSE) {
    const LevelDev<T> &Cc = mm.lev[l + 1];
    if (idx >= Cc.g.n[0] * Cc.g.n[1] * Cc.g.n[2]) return;
    unflat(Cc.g, idx, i, j, k);
    if (OP == EL_RESTRICT) {
      // the fine residuals at C-1, C, C+1 in place (k_resid_restrict's
      // expressions), so no separate residual phase
      int f[3] = { i, j, k };
      const int d = F.dir;
      f[d] *= 2;
      const long iF = LIDX(F.g, f[0], f[1], f[2]);
      const long sd = F.g.s[d];
      const T rC = resid_value<T, false>(F.g, F.A, F.x, F.b, f[0], f[1], f[2]);
      T rlo = T(0), rhi = T(0);
      if (f[d] > 0) {
        int q[3] = { f[0], f[1], f[2] };
        q[d] -= 1;
        rlo = F.p2[iF - sd] * resid_value<T, false>(F.g, F.A, F.x, F.b, q[0], q[1], q[2]);
      }
      if (f[d] + 1 < F.g.n[d]) {
        int q[3] = { f[0], f[1], f[2] };
        q[d] += 1;
        rhi = F.p1[iF + sd] * resid_value<T, false>(F.g, F.A, F.x, F.b, q[0], q[1], q[2]);
      }
      Cc.b[LIDX(Cc.g, i, j, k)] = rC + (rlo + rhi);
    }
    else
      coarse_point<T, false>(F.g, Cc.g, F.A, Cc.A, F.dir, F.o1, F.o2, F.p1, F.p2, i, j, k);
    return;
  }
  if (idx >= F.g.n[0] * F.g.n[1] * F.g.n[2]) return;
  unflat(F.g, idx, i, j, k);
  const long ip = LIDX(F.g, i, j, k);
  if (OP == EL_CONV_IN) F.b[ip] = (T)mm.bin[ip];
  else if (OP == EL_FILL) F.x[ip] = T(0);
  else if (OP == EL_PROLONG) {
    const LevelDev<T> &Cc = mm.lev[l + 1];
    prolong_point(F.g, Cc.g, F.dir, Cc.x, F.x, F.p1, F.p2, i, j, k);
  } else if (OP == EL_CONV_OUT) mm.xout[ip] = (double)F.x[ip];
}

// the one-workgroup part of every listed member: levels lc .. L-1, with the
// conversions of the right-hand side and the result when lc = 0
template <typename T>
__global__ __launch_bounds__(CWG) void k_ens_vcycle(const EnsLv<T> *__restrict__ M,
                                                    const int *__restrict__ list)
{
  const EnsLv<T> &mm = M[list[blockIdx.x]];
  if (!mm.act) return;                          // uniform over the workgroup
  const LevelDev<T> &L0 = mm.lev[0];
  if (mm.lc == 0) FOR_POINTS(L0.g, L0.b[LIDX(L0.g, i, j, k)] = (T)mm.bin[LIDX(L0.g, i, j, k)]);
  vcycle_block<T, false>(mm.lev, mm.lc, mm.L, mm.pre, mm.post, mm.symmetric);
  if (mm.lc == 0) FOR_POINTS(L0.g, mm.xout[LIDX(L0.g, i, j, k)] = (double)L0.x[LIDX(L0.g, i, j, k)]);
}

// the same with the member's levels lc .. L-1 in LDS (every member's
// bottom fits the plan's LDS budget)
template <typename T>
__global__ __launch_bounds__(CWG) void k_ens_vcycle_lds(const EnsLv<T> *__restrict__ M,
                                                        const int *__restrict__ list)
{
  const EnsLv<T> &mm = M[list[blockIdx.x]];
  if (!mm.act) return;                          // uniform over the workgroup
  vcycle_lds<T>(mm.lev, mm.lc, mm.L, mm.pre, mm.post, mm.symmetric,
                mm.lc == 0 ? mm.bin : nullptr, mm.lc == 0 ? mm.xout : nullptr);
}

template <typename T>
__global__ __launch_bounds__(CWG) void k_ens_mg_coarse(const EnsLv<T> *__restrict__ M,
                                                       const int *__restrict__ list)
{
  const EnsLv<T> &mm = M[list[blockIdx.x]];
  if (!mm.act) return;                          // uniform over the workgroup
  for (int l = mm.lc; l <= mm.L - 2; l++) {
    const LevelDev<T> &F = mm.lev[l], &Cc = mm.lev[l + 1];
    FOR_POINTS(Cc.g, (coarse_point<T, false>(F.g, Cc.g, F.A, Cc.A, F.dir, F.o1, F.o2, F.p1,
                                             F.p2, i, j, k)));
  }
}

extern "C" void pfhip_mg_level0(pfhip_mg *mg, void **A, int *prec)
{
  for (int q = 0; q < 4; q++) A[q] = mg->A[0][q];
  *prec = mg->prec;
}

static long ps_points(const pfhip_mg *mg, int l, int ps)
{
  if (ps == PS_COARSE) return npts(mg->n[l + 1]);
  if (ps == PS_HALF) return (long)((mg->n[l][0] + 1) / 2) * mg->n[l][1] * mg->n[l][2];
  return npts(mg->n[l]);
}

extern "C" long pfhip_ens_mg_plan_launches(const pfhip_ens_mgplan *p)
{
  return p ? p->launches : 0;
}

extern "C" int pfhip_ens_mg_plan_destroy(pfhip_ens_mgplan *p)
{
  if (!p) return PFHIP_OK;
  for (int l = 0; l < MAXL; l++)
    for (int s = 0; s < PS_N; s++) if (p->dseg[l][s]) pfhip_free(p->dseg[l][s]);
  for (int q = 0; q < 2; q++) if (p->gx[q]) hipGraphExecDestroy(p->gx[q]);
  if (p->cs) hipStreamDestroy(p->cs);
  if (p->dM) pfhip_free(p->dM);
  if (p->dlist) pfhip_free(p->dlist);
  free(p->lc);
  free(p);
  return PFHIP_OK;
}

extern "C" int pfhip_ens_mg_plan_create(pfhip_mg *const *mgs, int n, pfhip_ens_mgplan **out)
{
  pfhip_ens_mgplan *p = (pfhip_ens_mgplan *)calloc(1, sizeof(pfhip_ens_mgplan));
  int *order = (int *)malloc(sizeof(int) * (size_t)n);
  EnsLSeg *h = (EnsLSeg *)malloc(sizeof(EnsLSeg) * (size_t)n);
  if (p) p->lc = (int *)malloc(sizeof(int) * (size_t)n);
  if (!p || !order || !h || !p->lc) {
    free(order); free(h); pfhip_ens_mg_plan_destroy(p);
    pfhip_set_error("out of host memory");
    return PFHIP_ERR_NOMEM;
  }
  p->n = n;
  p->prec = mgs[0]->prec;
  p->pre = mgs[0]->pre; p->post = mgs[0]->post; p->symmetric = mgs[0]->symmetric;
  const char *ge = getenv("PFHIP_ENS_GRAPH");
  p->use_graph = (ge && *ge) ? atoi(ge) : 1;
  int rc = PFHIP_OK;
  const long small = ens_small();
  // the one-workgroup part in LDS: each member's first level from which
  // every coarser level fits the budget (no PFHIP_ENS_MG_SMALL given)
  const char *se = getenv("PFHIP_ENS_MG_SMALL");
  const char *fl = getenv("PFHIP_MG_LDS");
  const long lds_budget = (fl && *fl) ? atol(fl) : MG_LDS_DEFAULT;
  int dev = 0, lds_max = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess)
    lds_max = 0;
  const bool lds = !(se && *se) && lds_budget > 0 && lds_budget <= lds_max;
  const size_t es = (p->prec == 64) ? sizeof(double) : sizeof(float);
  const size_t rec = (p->prec == 64) ? sizeof(LevelDev<double>) : sizeof(LevelDev<float>);
  for (int m = 0; m < n; m++) {
    const pfhip_mg *mg = mgs[m];
    if (mg->prec != p->prec || mg->full || mg->pre != p->pre || mg->post != p->post ||
        mg->symmetric != p->symmetric) {
      pfhip_set_error("ensemble member %d: members need one multigrid precision and sweep "
                      "schedule and the PFSymmetric matrix", m);
      rc = PFHIP_ERR_ARG;
    }
    int lc = mg->L - 1;
    if (lds) {
      for (int l = 0; l < mg->L; l++) {
        const long b = lds_bottom_bytes(mg->n, l, mg->L, es, rec);
        if (b <= lds_budget) {
          lc = l;
          if (b > p->lds_bytes) p->lds_bytes = b;
          break;
        }
      }
    } else {
      for (int l = 0; l < mg->L; l++)
        if (npts(mg->n[l]) <= small) { lc = l; break; }
    }
    p->lc[m] = lc;
    if (lc > p->G) p->G = lc;
    order[m] = m;
  }
  if (rc == PFHIP_OK && p->lds_bytes > 0) {
    hipError_t e = (p->prec == 64)
        ? hipFuncSetAttribute((const void *)k_ens_vcycle_lds<double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes)
        : hipFuncSetAttribute((const void *)k_ens_vcycle_lds<float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_bytes);
    if (e != hipSuccess) {
      pfhip_set_error("hipFuncSetAttribute for the ensemble's LDS V-cycle: %s",
                      hipGetErrorString(e));
      rc = PFHIP_ERR_HIP;
    }
  }
  // members in decreasing size, so the largest segments start first
  for (int a = 1; a < n; a++) {
    int v = order[a], q = a;
    while (q > 0 && npts(mgs[order[q - 1]]->n[0]) < npts(mgs[v]->n[0])) {
      order[q] = order[q - 1];
      q--;
    }
    order[q] = v;
  }
  for (int l = 0; l < p->G && rc == PFHIP_OK; l++) {
    for (int s = 0; s < PS_N && rc == PFHIP_OK; s++) {
      int ns = 0, nb = 0;
      for (int q = 0; q < n; q++) {
        const int m = order[q];
        if (p->lc[m] <= l) continue;
        h[ns].m = m;
        h[ns].blk0 = nb;
        nb += (int)((ps_points(mgs[m], l, s) + ENS_LBLK - 1) / ENS_LBLK);
        ns++;
      }
      p->nseg[l][s] = ns;
      p->nblk[l][s] = nb;
      if (ns == 0) continue;
      rc = pfhip_malloc((void **)&p->dseg[l][s], (long)sizeof(EnsLSeg) * ns);
      if (rc == PFHIP_OK &&
          hipMemcpy(p->dseg[l][s], h, sizeof(EnsLSeg) * ns, hipMemcpyHostToDevice) != hipSuccess) {
        pfhip_set_error("hipMemcpy of an ensemble segment table");
        rc = PFHIP_ERR_HIP;
      }
    }
  }
  const size_t mb = (size_t)n * (p->prec == 32 ? sizeof(EnsLv<float>) : sizeof(EnsLv<double>));
  if (rc == PFHIP_OK) rc = pfhip_malloc(&p->dM, (long)mb);
  if (rc == PFHIP_OK) rc = pfhip_malloc((void **)&p->dlist, (long)sizeof(int) * n);
  if (rc == PFHIP_OK) {
    for (int m = 0; m < n; m++) order[m] = m;
    if (hipMemcpy(p->dlist, order, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) {
      pfhip_set_error("hipMemcpy of the ensemble member list");
      rc = PFHIP_ERR_HIP;
    }
  }
  if (rc == PFHIP_OK && p->use_graph &&
      hipStreamCreateWithFlags(&p->cs, hipStreamNonBlocking) != hipSuccess) {
    pfhip_set_error("hipStreamCreate for the ensemble graph capture");
    rc = PFHIP_ERR_HIP;
  }
  free(order);
  free(h);
  if (rc) { pfhip_ens_mg_plan_destroy(p); return rc; }
  *out = p;
  return PFHIP_OK;
}

template <typename T>
static void ens_phase(pfhip_ens_mgplan *p, int l, int ps, int op, int par, int init,
                      hipStream_t s)
{
  if (p->nblk[l][ps] == 0) return;
  p->launches++;
  hipLaunchKernelGGL(k_ens_lvl<T>, dim3(p->nblk[l][ps]), dim3(ENS_LBLK), 0, s, p->dseg[l][ps],
                     p->nseg[l][ps], (const EnsLv<T> *)p->dM, l, op, par, init);
}

template <typename T>
static void ens_sweeps(pfhip_ens_mgplan *p, int l, int zero, int max_iter, hipStream_t s)
{
  HalfSweep seq[2 * MAX_SWEEPS + 4];
  int fill;
  const int ns = sweep_list(p->symmetric, zero, max_iter, seq, &fill);
  if (fill) ens_phase<T>(p, l, PS_FINE, EL_FILL, 0, 0, s);
  for (int q = 0; q < ns; q++) ens_phase<T>(p, l, PS_HALF, EL_HALF, seq[q].par, seq[q].init, s);
}

// The launch sequence of one round: the coarse operators (apply = 0) or one
// V-cycle (apply = 1) of every member whose entry in the member table is
// active.  It reads only the plan's fixed tables, so it is the same in
// every round and can be captured once.
template <typename T>
static int ens_sequence(pfhip_ens_mgplan *p, int apply, hipStream_t s)
{
  const EnsLv<T> *dM = (const EnsLv<T> *)p->dM;
  if (!apply) {
    for (int l = 0; l < p->G; l++) {
      ens_phase<T>(p, l, PS_COARSE, EL_COARSE, 0, 0, s);
      LAUNCHCHK();
    }
    hipLaunchKernelGGL(k_ens_mg_coarse<T>, dim3(p->n), dim3(CWG), 0, s, dM, p->dlist);
    LAUNCHCHK();
    p->launches++;
    return PFHIP_OK;
  }
  if (p->G > 0) ens_phase<T>(p, 0, PS_FINE, EL_CONV_IN, 0, 0, s);
  for (int l = 0; l < p->G; l++) {
    ens_sweeps<T>(p, l, 1, p->pre, s);
    ens_phase<T>(p, l, PS_COARSE, EL_RESTRICT, 0, 0, s);
    LAUNCHCHK();
  }
  if (p->lds_bytes > 0)
    hipLaunchKernelGGL(k_ens_vcycle_lds<T>, dim3(p->n), dim3(CWG), (unsigned)p->lds_bytes, s,
                       dM, p->dlist);
  else
    hipLaunchKernelGGL(k_ens_vcycle<T>, dim3(p->n), dim3(CWG), 0, s, dM, p->dlist);
  LAUNCHCHK();
  p->launches++;
  for (int l = p->G - 1; l >= 0; l--) {
    ens_phase<T>(p, l, PS_FINE, EL_PROLONG, 0, 0, s);
    ens_sweeps<T>(p, l, 0, p->post, s);
    LAUNCHCHK();
  }
  if (p->G > 0) ens_phase<T>(p, 0, PS_FINE, EL_CONV_OUT, 0, 0, s);
  LAUNCHCHK();
  return PFHIP_OK;
}

// capture the round's launch sequence on the plan's stream (nothing runs
// during the capture) and instantiate it; kernel nodes only
template <typename T>
static int ens_capture(pfhip_ens_mgplan *p, int apply)
{
  hipGraph_t g = nullptr;
  const long l0 = p->launches;
  HIPCHK(hipStreamBeginCapture(p->cs, hipStreamCaptureModeRelaxed));
  int rc = ens_sequence<T>(p, apply, p->cs);
  hipError_t e = hipStreamEndCapture(p->cs, &g);
  p->glaunch[apply] = p->launches - l0;
  p->launches = l0;
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess || !g) {
    pfhip_set_error("ensemble graph capture: %s", hipGetErrorString(e));
    return PFHIP_ERR_HIP;
  }
  e = hipGraphInstantiate(&p->gx[apply], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    p->gx[apply] = nullptr;
    pfhip_set_error("ensemble graph instantiation: %s", hipGetErrorString(e));
    return PFHIP_ERR_HIP;
  }
  return PFHIP_OK;
}

template <typename T>
static int ens_mg_launch(pfhip_ens_mgplan *p, pfhip_ens_req *const *rq, int n, int apply,
                         pfhip_ens_stage *st, hipStream_t s)
{
  // the member table of this round: every member inactive but those with a
  // request; staged through the ring, then copied to the plan's fixed table
  const long mbytes = (long)sizeof(EnsLv<T>) * p->n;
  EnsLv<T> *M = (EnsLv<T> *)pfhip_ens_reserve(st, mbytes, s);
  if (!M) return PFHIP_ERR_HIP;
  memset(M, 0, (size_t)mbytes);
  for (int r = 0; r < n; r++) {
    pfhip_mg *mg = rq[r]->mg;
    const int m = rq[r]->member;
    if (m < 0 || m >= p->n) {
      pfhip_set_error("ensemble multigrid: request of member %d outside the plan", m);
      return PFHIP_ERR_ARG;
    }
    EnsLv<T> &e = M[m];
    e.lev = (const LevelDev<T> *)mg->levdev;
    e.bin = rq[r]->x;
    e.xout = rq[r]->z;
    e.L = mg->L; e.lc = p->lc[m];
    e.pre = mg->pre; e.post = mg->post; e.symmetric = mg->symmetric;
    e.act = 1;
  }
  void *dstage = pfhip_ens_commit(st, mbytes, s);
  if (!dstage) return PFHIP_ERR_HIP;
  HIPCHK(hipMemcpyAsync(p->dM, dstage, (size_t)mbytes, hipMemcpyDeviceToDevice, s));
  if (p->use_graph) {
    if (!p->gx[apply]) {
      int rc = ens_capture<T>(p, apply);
      if (rc) return rc;
    }
    HIPCHK(hipGraphLaunch(p->gx[apply], s));
    p->launches += p->glaunch[apply];
    return PFHIP_OK;
  }
  return ens_sequence<T>(p, apply, s);
}

static int ens_mg_exec(pfhip_ens_mgplan *p, pfhip_ens_req *const *rq, int n, int apply,
                       pfhip_ens_stage *st, void *stream)
{
  if (n <= 0) return PFHIP_OK;
  for (int r = 0; r < n; r++) {
    const pfhip_mg *mg = rq[r]->mg;
    if (mg->prec != p->prec || mg->full || mg->pre != p->pre || mg->post != p->post ||
        mg->symmetric != p->symmetric) {
      pfhip_set_error("ensemble members need one multigrid precision and sweep schedule and "
                      "the PFSymmetric preconditioner matrix");
      return PFHIP_ERR_ARG;
    }
  }
  return (p->prec == 32) ? ens_mg_launch<float>(p, rq, n, apply, st, STREAM(stream))
                         : ens_mg_launch<double>(p, rq, n, apply, st, STREAM(stream));
}

extern "C" int pfhip_ens_exec_mg_coarse(pfhip_ens_mgplan *p, pfhip_ens_req *const *rq, int n,
                                        pfhip_ens_stage *st, void *stream)
{
  return ens_mg_exec(p, rq, n, 0, st, stream);
}

extern "C" int pfhip_ens_exec_mg_apply(pfhip_ens_mgplan *p, pfhip_ens_req *const *rq, int n,
                                       pfhip_ens_stage *st, void *stream)
{
  return ens_mg_exec(p, rq, n, 1, st, stream);
}
