// core.hip -- libpfhip: C-ABI plumbing, N_Vector operations with
// deterministic wavefront-shuffle reductions, and the Richards physics
// kernels (van Genuchten closures, residual, Jacobian, 7-point matvec,
// kinematic overland flux) for gfx950.
//
// Layout: ParFlow Subvector layout, one ghost layer, x fastest.  Interior
// kernels launch 64x4 thread blocks (one wavefront per x-row segment, so
// every load along x is coalesced) with one z-plane per gridDim.z slice.
#include <hip/hip_runtime.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfhip_internal.h"
#include "device_util.h"
#include "pf_math.h"

// --------------------------------------------------------------------------
// errors
// --------------------------------------------------------------------------
static thread_local char g_err[512] = "";

extern "C" void pfhip_set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char *pfhip_last_error(void) { return g_err; }

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      pfhip_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x,             \
                      hipGetErrorString(e_));                              \
      return PFHIP_ERR_HIP;                                                \
    }                                                                      \
  } while (0)

#define LAUNCHCHK()                                                        \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess) {                                                \
      pfhip_set_error("%s:%d kernel launch: %s", __FILE__, __LINE__,      \
                      hipGetErrorString(e_));                              \
      return PFHIP_ERR_HIP;                                                \
    }                                                                      \
  } while (0)

#define STREAM(s) ((hipStream_t)(s))

// --------------------------------------------------------------------------
// live timing with HIP events (pfhip_prof_*): when a timer id is enabled,
// every launch of that kernel (or every call of that solver phase) is
// bracketed by an event pair on the stream it is launched on.
// --------------------------------------------------------------------------
struct ProfState {
  int enabled[PFHIP_PROF_NKERNELS];
  int cap[PFHIP_PROF_NKERNELS];
  int count[PFHIP_PROF_NKERNELS];
  hipEvent_t *ev[PFHIP_PROF_NKERNELS];
};
static ProfState g_prof;

extern "C" int pfhip_prof_enable(int kernel_id, int max_launches)
{
  if (kernel_id < 0 || kernel_id >= PFHIP_PROF_NKERNELS || max_launches < 1) {
    pfhip_set_error("prof_enable: bad kernel id %d", kernel_id);
    return PFHIP_ERR_ARG;
  }
  g_prof.enabled[kernel_id] = 0;
  if (g_prof.ev[kernel_id]) {
    for (int i = 0; i < 2 * g_prof.cap[kernel_id]; i++) hipEventDestroy(g_prof.ev[kernel_id][i]);
    free(g_prof.ev[kernel_id]);
    g_prof.ev[kernel_id] = nullptr;
    g_prof.cap[kernel_id] = 0;
  }
  hipEvent_t *ev = (hipEvent_t *)calloc(2 * (size_t)max_launches, sizeof(hipEvent_t));
  if (!ev) { pfhip_set_error("prof_enable: out of host memory"); return PFHIP_ERR_NOMEM; }
  for (int i = 0; i < 2 * max_launches; i++) {
    hipError_t e = hipEventCreate(&ev[i]);
    if (e != hipSuccess) {
      for (int q = 0; q < i; q++) hipEventDestroy(ev[q]);
      free(ev);
      pfhip_set_error("prof_enable: hipEventCreate: %s", hipGetErrorString(e));
      return PFHIP_ERR_HIP;
    }
  }
  g_prof.ev[kernel_id] = ev;
  g_prof.cap[kernel_id] = max_launches;
  g_prof.count[kernel_id] = 0;
  g_prof.enabled[kernel_id] = 1;
  return PFHIP_OK;
}

extern "C" int pfhip_prof_reset(int kernel_id)
{
  if (kernel_id < 0 || kernel_id >= PFHIP_PROF_NKERNELS) return PFHIP_ERR_ARG;
  g_prof.count[kernel_id] = 0;
  return PFHIP_OK;
}

extern "C" int pfhip_prof_read(int kernel_id, double *total_ms, long *launches)
{
  if (kernel_id < 0 || kernel_id >= PFHIP_PROF_NKERNELS) return PFHIP_ERR_ARG;
  double tot = 0.0;
  int n = g_prof.count[kernel_id];
  if (n > g_prof.cap[kernel_id]) n = g_prof.cap[kernel_id];
  for (int i = 0; i < n; i++) {
    float ms = 0.f;
    HIPCHK(hipEventSynchronize(g_prof.ev[kernel_id][2 * i + 1]));
    HIPCHK(hipEventElapsedTime(&ms, g_prof.ev[kernel_id][2 * i],
                               g_prof.ev[kernel_id][2 * i + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = n;
  return PFHIP_OK;
}

extern "C" int pfhip_prof_disable(int kernel_id)
{
  if (kernel_id < 0 || kernel_id >= PFHIP_PROF_NKERNELS) return PFHIP_ERR_ARG;
  g_prof.enabled[kernel_id] = 0;
  return PFHIP_OK;
}

extern "C" int pfhip_prof_enabled(int kernel_id)
{
  return (kernel_id >= 0 && kernel_id < PFHIP_PROF_NKERNELS) ? g_prof.enabled[kernel_id] : 0;
}

extern "C" void pfhip_prof_mark(int kernel_id, int end, void *stream)
{
  if (!g_prof.enabled[kernel_id]) return;
  int c = g_prof.count[kernel_id];
  if (c >= g_prof.cap[kernel_id]) return;
  hipEventRecord(g_prof.ev[kernel_id][2 * c + end], STREAM(stream));
  if (end) g_prof.count[kernel_id] = c + 1;
}
#define PROF_BEGIN(id, st) pfhip_prof_mark((id), 0, (st))
#define PROF_END(id, st) pfhip_prof_mark((id), 1, (st))

// --------------------------------------------------------------------------
// memory
// --------------------------------------------------------------------------
extern "C" int pfhip_device_count(int *count)
{
  HIPCHK(hipGetDeviceCount(count));
  return PFHIP_OK;
}

extern "C" int pfhip_malloc(void **ptr, long bytes)
{
  if (bytes <= 0) { *ptr = nullptr; return PFHIP_OK; }
  hipError_t e = hipMalloc(ptr, (size_t)bytes);
  if (e != hipSuccess) {
    pfhip_set_error("hipMalloc(%ld): %s", bytes, hipGetErrorString(e));
    return PFHIP_ERR_NOMEM;
  }
  return PFHIP_OK;
}

extern "C" int pfhip_free(void *ptr)
{
  if (ptr) HIPCHK(hipFree(ptr));
  return PFHIP_OK;
}

extern "C" int pfhip_memcpy_h2d(void *dst, const void *src, long bytes)
{
  HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return PFHIP_OK;
}

extern "C" int pfhip_memcpy_d2h(void *dst, const void *src, long bytes)
{
  HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return PFHIP_OK;
}

extern "C" int pfhip_memset(void *dst, int value, long bytes)
{
  HIPCHK(hipMemset(dst, value, (size_t)bytes));
  return PFHIP_OK;
}

extern "C" int pfhip_synchronize(void)
{
  HIPCHK(hipDeviceSynchronize());
  return PFHIP_OK;
}

extern "C" int pfhip_dev_alloc(double **p, long n)
{
  return pfhip_malloc((void **)p, n * (long)sizeof(double));
}

__global__ void k_fill(double *__restrict__ p, long n, double v)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

extern "C" int pfhip_zero(double *p, long n, void *stream)
{
  if (n <= 0) return PFHIP_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0,
                     STREAM(stream), p, n, 0.0);
  LAUNCHCHK();
  return PFHIP_OK;
}

// --------------------------------------------------------------------------
// geometry helpers
// --------------------------------------------------------------------------
struct Geom {
  int nx, ny, nz, sy, sz;
};

static inline Geom geom_of(const pfhip_problem *pb)
{
  Geom g;
  g.nx = pb->d.nx; g.ny = pb->d.ny; g.nz = pb->d.nz;
  g.sy = pb->sy; g.sz = pb->sz;
  return g;
}

static inline dim3 grid3(const Geom &g)
{
  return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
}
#define BLOCK3 dim3(64, 4, 1)

// --------------------------------------------------------------------------
// problem
// --------------------------------------------------------------------------
static long res_tiles(const Geom &g);
static long res_bc_blocks(const Geom &g);

static int face_ncells(const pfhip_problem_desc *d, int f)
{
  if (f < 2) return d->ny * d->nz;
  if (f < 4) return d->nx * d->nz;
  return d->nx * d->ny;
}

// VanGComputeTable (PL/problem_phase_rel_perm.c:116-262) on the host: kr and
// its derivative formula at points+1 heads spaced |min_head| / (points-1)
// from 0, the Fritsch-Carlson monotone Hermite slopes of both (Fritsch &
// Carlson, SIAM J. Numer. Anal. 17 (1980)) and the linear slopes, stored
// x | a | d | a_der | d_der | slope | slope_der
static void build_kr_table(int method, int points, double min_head, double alpha, double n,
                           double *tab)
{
  const int N = points, N1 = points + 1;
  double *x = tab, *a = tab + N1, *d = tab + 2 * N1, *a_der = tab + 3 * N1;
  double *d_der = tab + 4 * N1, *slope = tab + 5 * N1, *slope_der = tab + 6 * N1;
  const double step = min_head / (double)(N - 1), astep = fabs(step);
  const double m = 1.0e0 - (1.0e0 / n);
  for (int i = 0; i < N1; i++) {
    x[i] = fabs(i * step);
    double opahn = 1.0 + pow(alpha * x[i], n);
    double ahnm1 = pow(alpha * x[i], n - 1);
    a[i] = pow(1.0 - ahnm1 / pow(opahn, m), 2) / pow(opahn, m / 2);
    double coeff = 1.0 - ahnm1 * pow(opahn, -m);
    a_der[i] = 2.0 * (coeff / pow(opahn, m / 2)) *
                   ((n - 1) * pow(alpha * x[i], n - 2) * alpha * pow(opahn, -m) -
                    ahnm1 * m * pow(opahn, -(m + 1)) * n * alpha * ahnm1) +
               pow(coeff, 2) * (m / 2) * pow(opahn, -(m + 2) / 2) * n * alpha * ahnm1;
    if (n < 2 && i == 0) a_der[i] = 0;   // unbounded at head 0 for 1 < n < 2
  }
  if (method == 1)
    for (int i = 0; i < N; i++) {
      slope[i] = (a[i + 1] - a[i]) / astep;
      slope_der[i] = (a_der[i + 1] - a_der[i]) / astep;
    }
  double *del = (double *)malloc(sizeof(double) * 2 * N1), *del_der = del + N1;
  for (int i = 0; i < N; i++) {
    double h = x[i + 1] - x[i];
    del[i] = (a[i + 1] - a[i]) / h;
    del_der[i] = (a_der[i + 1] - a_der[i]) / h;
  }
  d[0] = del[0]; d[N] = del[N - 1];
  d_der[0] = del_der[0]; d_der[N] = del_der[N - 1];
  for (int i = 1; i < N; i++) {
    d[i] = (del[i - 1] + del[i]) / 2;
    d_der[i] = (del_der[i - 1] + del_der[i]) / 2;
  }
  // limit the slopes interval by interval so the interpolant stays monotone
  auto limit = [](double *dd, const double *dl, int i) {
    if (dl[i] == 0.0) { dd[i] = 0; dd[i + 1] = 0; return; }
    double al = dd[i] / dl[i], be = dd[i + 1] / dl[i];
    double magn = pow(al, 2) + pow(be, 2);
    if (magn > 9.0) {
      dd[i] = 3 * al * dl[i] / magn;
      dd[i + 1] = 3 * be * dl[i] / magn;
    }
  };
  for (int i = 0; i < N; i++) {
    limit(d, del, i);
    limit(d_der, del_der, i);
  }
  free(del);
}

static double *noalloc_scratch();

extern "C" int pfhip_problem_create(const pfhip_problem_desc *desc,
                                    pfhip_problem **out)
{
  if (!desc || !out) { pfhip_set_error("null argument"); return PFHIP_ERR_ARG; }
  const pfhip_problem_desc *d = desc;
  if (d->nx < 1 || d->ny < 1 || d->nz < 1 || d->dx <= 0 || d->dy <= 0 ||
      d->dz <= 0) {
    pfhip_set_error("bad grid %dx%dx%d", d->nx, d->ny, d->nz);
    return PFHIP_ERR_ARG;
  }
  if (!d->kx || !d->ky || !d->kz || !d->poro || !d->ss || !d->alpha ||
      !d->nvg) {
    pfhip_set_error("problem fields must be device arrays");
    return PFHIP_ERR_ARG;
  }
  for (int f = 0; f < 6; f++) {
    const pfhip_bc_spec *b = &d->bc[f];
    if (b->type < 0 || b->type > 2 || (b->type == PFHIP_BC_OVERLAND && f != 5)) {
      pfhip_set_error("bad BC type %d on face %d", b->type, f);
      return PFHIP_ERR_ARG;
    }
    if (b->n_intervals < 1 || b->n_intervals > PFHIP_MAX_INTERVALS) {
      pfhip_set_error("bad BC cycle on face %d", f);
      return PFHIP_ERR_ARG;
    }
  }
  if (d->bc[5].type == PFHIP_BC_OVERLAND &&
      (!d->slope_x || !d->slope_y || !d->mann)) {
    pfhip_set_error("OverlandKinematic needs slope_x, slope_y, mann");
    return PFHIP_ERR_ARG;
  }
  if (d->vg_table_points != 0 &&
      (d->vg_table_points < 2 || !(d->vg_table_min_head < 0.0) || d->vg_table_method < 0 ||
       d->vg_table_method > 1 || !(d->vg_table_n > 1.0) || !(d->vg_table_alpha > 0.0))) {
    pfhip_set_error("bad van Genuchten table: points %d, min head %g, method %d",
                    d->vg_table_points, d->vg_table_min_head, d->vg_table_method);
    return PFHIP_ERR_ARG;
  }
  long ng = (long)(d->nx + 2) * (d->ny + 2) * (d->nz + 2);
  // the residual kernels load through 32-bit byte offsets from each array's
  // base (ldb): an FP64 ghosted array must stay below 4 GiB, 2^29 cells
  if (ng >= PFHIP_MAX_GHOSTED_CELLS) {
    pfhip_set_error("ghosted grid of %ld cells exceeds the limit of %ld (4 GiB per FP64 array)",
                    ng, (long)PFHIP_MAX_GHOSTED_CELLS);
    return PFHIP_ERR_ARG;
  }
  pfhip_problem *pb = (pfhip_problem *)calloc(1, sizeof(pfhip_problem));
  pb->d = *d;
  pb->sy = d->nx + 2;
  pb->sz = (d->nx + 2) * (d->ny + 2);
  pb->ng = ng;
  pb->n2 = (long)(d->nx + 2) * (d->ny + 2);
  pb->ncell = (long)d->nx * d->ny * d->nz;
  pb->overland = (d->bc[5].type == PFHIP_BC_OVERLAND);
  int rc = PFHIP_OK;
  for (int f = 0; f < 6 && rc == PFHIP_OK; f++) {
    int n = face_ncells(d, f);
    rc = pfhip_dev_alloc(&pb->bcv[f], n);
    if (rc == PFHIP_OK) {
      hipError_t e = hipHostMalloc((void **)&pb->h_bcv[f], sizeof(double) * n);
      if (e != hipSuccess) { pfhip_set_error("hipHostMalloc"); rc = PFHIP_ERR_NOMEM; }
    }
  }
  if (rc == PFHIP_OK) rc = pfhip_dev_alloc(&pb->krho, ng);
  if (rc == PFHIP_OK) rc = pfhip_dev_alloc(&pb->dkrho, ng);
  for (int q = 0; q < 4 && rc == PFHIP_OK; q++) rc = pfhip_dev_alloc(&pb->q2[q], pb->n2);
  if (rc == PFHIP_OK) rc = pfhip_dev_alloc(&pb->red_partial, PFHIP_RED_MAX_BLOCKS * PFHIP_MAX_MULTI);
  if (rc == PFHIP_OK) rc = pfhip_dev_alloc(&pb->red_dev, PFHIP_RED_SLOTS);
  if (rc == PFHIP_OK) {
    Geom g0 = geom_of(pb);
    pb->stat_nb = res_tiles(g0) + res_bc_blocks(g0);
    rc = pfhip_dev_alloc(&pb->stat_partial, 2 * pb->stat_nb);
  }
  if (rc == PFHIP_OK) {
    hipError_t e = hipHostMalloc((void **)&pb->h_red, sizeof(double) * PFHIP_MAX_MULTI);
    if (e != hipSuccess) { pfhip_set_error("hipHostMalloc"); rc = PFHIP_ERR_NOMEM; }
  }
  if (rc == PFHIP_OK) rc = pfhip_zero(pb->krho, ng, nullptr);
  if (rc == PFHIP_OK) rc = pfhip_zero(pb->dkrho, ng, nullptr);
  if (rc == PFHIP_OK && d->vg_table_points > 0) {
    const int N1 = d->vg_table_points + 1;
    double *host = (double *)calloc(7 * (size_t)N1, sizeof(double));
    if (!host) rc = PFHIP_ERR_NOMEM;
    if (rc == PFHIP_OK) {
      build_kr_table(d->vg_table_method, d->vg_table_points, d->vg_table_min_head,
                     d->vg_table_alpha, d->vg_table_n, host);
      rc = pfhip_dev_alloc(&pb->krtab, 7 * (long)N1);
    }
    if (rc == PFHIP_OK && hipMemcpy(pb->krtab, host, sizeof(double) * 7 * N1,
                                    hipMemcpyHostToDevice) != hipSuccess) {
      pfhip_set_error("hipMemcpy of the kr table");
      rc = PFHIP_ERR_HIP;
    }
    free(host);
  }
  // set-up of the no-allocation entry point: this device's scratch address
  if (rc == PFHIP_OK && !noalloc_scratch()) rc = PFHIP_ERR_HIP;
  if (rc != PFHIP_OK) { pfhip_problem_destroy(pb); return rc; }
  HIPCHK(hipDeviceSynchronize());
  *out = pb;
  return PFHIP_OK;
}

extern "C" int pfhip_problem_destroy(pfhip_problem *pb)
{
  if (!pb) return PFHIP_OK;
  for (int f = 0; f < 6; f++) {
    if (pb->bcv[f]) hipFree(pb->bcv[f]);
    if (pb->h_bcv[f]) hipHostFree(pb->h_bcv[f]);
  }
  if (pb->krho) hipFree(pb->krho);
  if (pb->dkrho) hipFree(pb->dkrho);
  for (int q = 0; q < 4; q++) if (pb->q2[q]) hipFree(pb->q2[q]);
  if (pb->red_partial) hipFree(pb->red_partial);
  if (pb->red_dev) hipFree(pb->red_dev);
  if (pb->stat_partial) hipFree(pb->stat_partial);
  if (pb->h_red) hipHostFree(pb->h_red);
  if (pb->krtab) hipFree(pb->krtab);
  free(pb);
  return PFHIP_OK;
}

// TimeCycleDataComputeIntervalNumber, PL/time_cycle_data.c:152-197
static int interval_number(const pfhip_problem_desc *d, const pfhip_bc_spec *b,
                           double time)
{
  int cycle = 0;
  for (int i = 0; i < b->n_intervals; i++) cycle += b->lengths[i];
  double bu = d->base_unit;
  if (b->repeat < 0 || time < (double)(b->repeat * cycle * bu)) {
    int dt_ = (int)lround(time / (bu / 1000)) / 1000;
    int ds_ = (int)lround(d->start_time / (bu / 1000)) / 1000;
    int completed = (dt_ - ds_ - 1) % cycle;
    int iv = 0, total = b->lengths[0];
    while (completed >= total) { iv++; total += b->lengths[iv]; }
    return iv;
  }
  return b->n_intervals - 1;
}

// BCPressure: PL/problem_bc_pressure.c:143-290 (DirEquilRefPatch with a
// constant density reduces to p = p_ref - rho g (z - z_ref), z = centre for
// side faces, face elevation for z faces), :582-612, :1070-1098.
// the BC values of face f at `time` into h (face_ncells values)
static void bc_face_values(const pfhip_problem *pb, int f, double time, double *h)
{
  const pfhip_problem_desc *d = &pb->d;
  const pfhip_bc_spec *b = &d->bc[f];
  int n = face_ncells(d, f);
  double val = b->values[interval_number(d, b, time)];
  {
    if (b->type != PFHIP_BC_DIR_EQUIL) {
      for (int c = 0; c < n; c++) h[c] = val;
    } else {
      for (int c = 0; c < n; c++) {
        int k;
        if (f < 2) k = c / d->ny;
        else if (f < 4) k = c / d->nx;
        else k = (f == 4) ? 0 : d->nz - 1;
        double z = d->z0 + (k + 0.5) * d->dz;
        if (f == 4) z = z - 0.5 * d->dz;
        if (f == 5) z = z + 0.5 * d->dz;
        double gneg = -d->gravity;
        double fcn_val = val - val - 0.5 * (d->rho + d->rho) * gneg * (z - b->z_ref);
        double pv = val - fcn_val / 1.0;
        h[c] = pv;
      }
    }
  }
}

extern "C" int pfhip_bc_pressure_async(pfhip_problem *pb, double time, int sync, void *stream)
{
  for (int f = 0; f < 6; f++) {
    bc_face_values(pb, f, time, pb->h_bcv[f]);
    HIPCHK(hipMemcpyAsync(pb->bcv[f], pb->h_bcv[f], sizeof(double) * face_ncells(&pb->d, f),
                          hipMemcpyHostToDevice, STREAM(stream)));
  }
  if (sync) HIPCHK(hipStreamSynchronize(STREAM(stream)));
  return PFHIP_OK;
}

extern "C" int pfhip_bc_pressure(pfhip_problem *pb, double time, void *stream)
{
  return pfhip_bc_pressure_async(pb, time, 1, stream);
}

// --------------------------------------------------------------------------
// deterministic reductions over interior rows
// --------------------------------------------------------------------------
// A block owns whole (j,k) rows, threads stride along x: one division per
// row, coalesced loads.  Block partials -> one final block, fixed order.

static inline int red_blocks(const Geom &g)
{
  long rows = (long)g.ny * g.nz;
  return (int)(rows < PFHIP_RED_MAX_BLOCKS ? rows : PFHIP_RED_MAX_BLOCKS);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// 256-thread block reduction; result valid in thread 0
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v)
{
  __shared__ double sh[4];
  v = MAX ? wave_max(v) : wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int i = 1; i < 4; i++) r = MAX ? fmax(r, sh[i]) : r + sh[i];
  }
  __syncthreads();
  return r;
}

enum RedOp { RED_DOT = 0, RED_MAXABS = 1, RED_L1 = 2, RED_SCSTEP = 3, RED_DQ = 4,
             RED_WL2 = 5, RED_NEGMAX = 6, RED_CONSTR = 7, RED_DOTSC = 8, RED_KINIT = 9 };

template <int OP>
__global__ __launch_bounds__(256) void k_reduce(Geom g, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                double *__restrict__ partial, int nb)
{
  double a0 = (OP == RED_NEGMAX) ? -INFINITY : 0.0, a1 = 0.0, a2 = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      if (OP == RED_DOT) a0 += x[ip] * y[ip];
      else if (OP == RED_WL2) { double pr = x[ip] * y[ip]; a0 += pr * pr; }
      else if (OP == RED_NEGMAX) a0 = fmax(a0, -x[ip]);
      else if (OP == RED_CONSTR) a0 += (y[ip] != 0.0 && x[ip] * y[ip] <= 0.0) ? 1.0 : 0.0;
      else if (OP == RED_MAXABS) a0 = fmax(a0, fabs(x[ip]));
      else if (OP == RED_L1) a0 += fabs(x[ip]);
      else if (OP == RED_SCSTEP) a0 = fmax(a0, fabs(y[ip] / (1.0 + fabs(x[ip]))));
      else if (OP == RED_KINIT) {
        // max |y| (RED_MAXABS of y), y.y and x.x (RED_DOT's products)
        const double fv = y[ip], pv = x[ip];
        a0 = fmax(a0, fabs(fv));
        a1 += fv * fv;
        a2 += pv * pv;
      } else if (OP == RED_DOTSC) {
        // y.y (RED_DOT's product with x = y) and RED_SCSTEP's maximum
        const double pv = y[ip];
        a0 += pv * pv;
        a1 = fmax(a1, fabs(pv / (1.0 + fabs(x[ip]))));
      } else {  // RED_DQ: u.v, v.v, |v|_1
        double u = x[ip], v = y[ip];
        a0 += u * v;
        a1 += v * v;
        a2 += fabs(v);
      }
    }
  }
  const bool MX = (OP == RED_MAXABS || OP == RED_SCSTEP || OP == RED_NEGMAX || OP == RED_KINIT);
  a0 = block_reduce<MX>(a0);
  if (OP == RED_DQ) { a1 = block_reduce<false>(a1); a2 = block_reduce<false>(a2); }
  if (OP == RED_DOTSC) a1 = block_reduce<true>(a1);
  if (OP == RED_KINIT) { a1 = block_reduce<false>(a1); a2 = block_reduce<false>(a2); }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = a0;
    if (OP == RED_DQ || OP == RED_KINIT) {
      partial[nb + blockIdx.x] = a1;
      partial[2 * nb + blockIdx.x] = a2;
    }
    if (OP == RED_DOTSC) partial[nb + blockIdx.x] = a1;
  }
}

// final fixed-order reduction of nv vectors of nb partials each
template <bool MAX>
__global__ __launch_bounds__(256) void k_reduce_final(const double *__restrict__ partial,
                                                      int nb, int nv,
                                                      double *__restrict__ out)
{
  for (int v = 0; v < nv; v++) {
    double a = MAX ? -INFINITY : 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      double t = partial[v * nb + i];
      a = MAX ? fmax(a, t) : a + t;
    }
    a = block_reduce<MAX>(a);
    if (threadIdx.x == 0) out[v] = a;
  }
}

static int fetch_results(pfhip_problem *pb, int nv, void *stream)
{
  HIPCHK(hipMemcpyAsync(pb->h_red, pb->red_dev, sizeof(double) * nv,
                        hipMemcpyDeviceToHost, STREAM(stream)));
  HIPCHK(hipStreamSynchronize(STREAM(stream)));
  return PFHIP_OK;
}

template <int OP>
static int run_reduce(pfhip_problem *pb, const double *x, const double *y,
                      int nv, void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_reduce<OP>, dim3(nb), dim3(256), 0, STREAM(stream), g, x,
                     y, pb->red_partial, nb);
  LAUNCHCHK();
  const bool MX = (OP == RED_MAXABS || OP == RED_SCSTEP || OP == RED_NEGMAX);
  hipLaunchKernelGGL(k_reduce_final<MX>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, nv, pb->red_dev);
  LAUNCHCHK();
  return fetch_results(pb, nv, stream);
}

extern "C" int pfhip_vec_dot(pfhip_problem *pb, const double *x, const double *y,
                             double *result, void *stream)
{
  int rc = run_reduce<RED_DOT>(pb, x, y, 1, stream);
  if (rc) return rc;
  *result = pb->h_red[0];
  return PFHIP_OK;
}

extern "C" int pfhip_vec_max_norm(pfhip_problem *pb, const double *x,
                                  double *result, void *stream)
{
  int rc = run_reduce<RED_MAXABS>(pb, x, x, 1, stream);
  if (rc) return rc;
  *result = pb->h_red[0];
  return PFHIP_OK;
}

extern "C" int pfhip_vec_l1_norm(pfhip_problem *pb, const double *x,
                                 double *result, void *stream)
{
  int rc = run_reduce<RED_L1>(pb, x, x, 1, stream);
  if (rc) return rc;
  *result = pb->h_red[0];
  return PFHIP_OK;
}

extern "C" int pfhip_vec_scsteplength(pfhip_problem *pb, const double *u,
                                      const double *s, double *result, void *stream)
{
  int rc = run_reduce<RED_SCSTEP>(pb, u, s, 1, stream);
  if (rc) return rc;
  *result = pb->h_red[0];
  return PFHIP_OK;
}

// p.p and max |p| / (1 + |u|) from one pass: the same values as
// pfhip_vec_dot(p, p) and pfhip_vec_scsteplength(u, p), one synchronisation
int pfhip_vec_dot_scsteplength(pfhip_problem *pb, const double *u, const double *p,
                               double *pp, double *rlength, void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_reduce<RED_DOTSC>, dim3(nb), dim3(256), 0, STREAM(stream), g, u, p,
                     pb->red_partial, nb);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 1, pb->red_dev);
  hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial + nb, nb, 1, pb->red_dev + 1);
  LAUNCHCHK();
  int rc = fetch_results(pb, 2, stream);
  if (rc) return rc;
  *pp = pb->h_red[0];
  *rlength = pb->h_red[1];
  return PFHIP_OK;
}

// KINSOL's start (KS/kinsol.c:690-760): max |F| for the initial stop test,
// F.F and u.u from one pass, with the values of pfhip_vec_max_norm(F),
// pfhip_vec_dot(F, F) and pfhip_vec_dot(u, u) and one synchronisation
int pfhip_vec_kinsol_init_norms(pfhip_problem *pb, const double *u, const double *f,
                                double *fmax, double *ff, double *uu, void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_reduce<RED_KINIT>, dim3(nb), dim3(256), 0, STREAM(stream), g, u, f,
                     pb->red_partial, nb);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 1, pb->red_dev);
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial + nb, nb, 2, pb->red_dev + 1);
  LAUNCHCHK();
  int rc = fetch_results(pb, 3, stream);
  if (rc) return rc;
  *fmax = pb->h_red[0];
  *ff = pb->h_red[1];
  *uu = pb->h_red[2];
  return PFHIP_OK;
}

extern "C" int pfhip_vec_dq_sums(pfhip_problem *pb, const double *u,
                                 const double *v, double *sutsv, double *vtv,
                                 double *l1, void *stream)
{
  int rc = run_reduce<RED_DQ>(pb, u, v, 3, stream);
  if (rc) return rc;
  *sutsv = pb->h_red[0];
  *vtv = pb->h_red[1];
  *l1 = pb->h_red[2];
  return PFHIP_OK;
}

extern "C" int pfhip_vec_wl2_norm(pfhip_problem *pb, const double *x, const double *w,
                                  double *result, void *stream)
{
  int rc = run_reduce<RED_WL2>(pb, x, w, 1, stream);
  if (rc) return rc;
  *result = sqrt(pb->h_red[0]);
  return PFHIP_OK;
}

extern "C" int pfhip_vec_wrms_norm(pfhip_problem *pb, const double *x, const double *w,
                                   double *result, void *stream)
{
  int rc = run_reduce<RED_WL2>(pb, x, w, 1, stream);
  if (rc) return rc;
  *result = sqrt(pb->h_red[0] / (double)pb->ncell);
  return PFHIP_OK;
}

extern "C" int pfhip_vec_min(pfhip_problem *pb, const double *x, double *result, void *stream)
{
  int rc = run_reduce<RED_NEGMAX>(pb, x, x, 1, stream);
  if (rc) return rc;
  *result = -pb->h_red[0];
  return PFHIP_OK;
}

extern "C" int pfhip_vec_constr_prod_pos(pfhip_problem *pb, const double *c, const double *x,
                                         int *result, void *stream)
{
  int rc = run_reduce<RED_CONSTR>(pb, x, c, 1, stream);
  if (rc) return rc;
  *result = (pb->h_red[0] == 0.0) ? 1 : 0;
  return PFHIP_OK;
}

__global__ void k_inv_test(Geom g, const double *__restrict__ x, double *__restrict__ z,
                           double *__restrict__ partial);

extern "C" int pfhip_vec_inv_test(pfhip_problem *pb, const double *x, double *z, int *result,
                                  void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_inv_test, dim3(nb), dim3(256), 0, STREAM(stream), g, x, z,
                     pb->red_partial);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 1, pb->red_dev);
  LAUNCHCHK();
  int rc = fetch_results(pb, 1, stream);
  if (rc) return rc;
  *result = (pb->h_red[0] == 0.0) ? 1 : 0;
  return PFHIP_OK;
}

// KINStop's two tests in one pass (KS/kinsol.c:1530-1550 with uscale =
// fscale = 1): max |fval| and max |unew - uu| / (1 + |unew|), the step being
// formed on the fly with the rounding of the separate difference pass.
__global__ __launch_bounds__(256) void k_stop_red(Geom g, const double *__restrict__ fval,
                                                  const double *__restrict__ unew,
                                                  const double *__restrict__ uu,
                                                  double *__restrict__ partial, int nb)
{
  double a0 = 0.0, a1 = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      a0 = fmax(a0, fabs(fval[ip]));
      double un = unew[ip];
      double st = 1.0 * un + (-1.0) * uu[ip];
      a1 = fmax(a1, fabs(st / (1.0 + fabs(un))));
    }
  }
  a0 = block_reduce<true>(a0);
  a1 = block_reduce<true>(a1);
  if (threadIdx.x == 0) { partial[blockIdx.x] = a0; partial[nb + blockIdx.x] = a1; }
}

int pfhip_kin_stop_norms(pfhip_problem *pb, const double *fval, const double *unew,
                         const double *uu, double *fmax, double *steplength, void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_stop_red, dim3(nb), dim3(256), 0, STREAM(stream), g, fval, unew, uu,
                     pb->red_partial, nb);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 2, pb->red_dev);
  LAUNCHCHK();
  int rc = fetch_results(pb, 2, stream);
  if (rc) return rc;
  *fmax = pb->h_red[0];
  *steplength = pb->h_red[1];
  return PFHIP_OK;
}

// ---- water-balance storage ---------------------------------------------
// calculate_subsurface_storage / calculate_surface_storage
// (pftools/python/parflow/tools/hydrology.py:136-205) summed over the box,
// with their per-cell operation order; row-owning blocks and the fixed-order
// final pass make the sums reproducible bit for bit.
__global__ __launch_bounds__(256) void k_storage(Geom g, double dx, double dy, double dz,
                                                 const double *__restrict__ p,
                                                 const double *__restrict__ s,
                                                 const double *__restrict__ poro,
                                                 const double *__restrict__ ss,
                                                 double *__restrict__ partial, int nb)
{
  double sub = 0.0, sur = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      double pv = p[ip], sv = s[ip];
      double inc = poro[ip] * sv * dz * dx * dy;
      double comp = pv * sv * ss[ip] * dz * dx * dy;
      if (pv < 0.0) comp = 0.0;
      sub += inc + comp;
      if (k == g.nz - 1) {
        double t = pv * dx * dy;
        sur += (t < 0.0) ? 0.0 : t;
      }
    }
  }
  sub = block_reduce<false>(sub);
  sur = block_reduce<false>(sur);
  if (threadIdx.x == 0) { partial[blockIdx.x] = sub; partial[nb + blockIdx.x] = sur; }
}

extern "C" int pfhip_water_balance_storage(pfhip_problem *pb, const double *p,
                                           const double *s, double *subsurface,
                                           double *surface, void *stream)
{
  if (!p || !s || !subsurface || !surface) {
    pfhip_set_error("water_balance_storage: NULL argument");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_storage, dim3(nb), dim3(256), 0, STREAM(stream), g, pb->d.dx,
                     pb->d.dy, pb->d.dz, p, s, pb->d.poro, pb->d.ss, pb->red_partial, nb);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 2, pb->red_dev);
  LAUNCHCHK();
  int rc = fetch_results(pb, 2, stream);
  if (rc) return rc;
  *subsurface = pb->h_red[0];
  *surface = pb->h_red[1];
  return PFHIP_OK;
}

// ---- the same storage sums without allocating ------------------------------
// For callers that may not allocate device memory: the block partials and
// the two results live in a file-scope device array, one copy per GPU, held
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): the address of
// each GPU's copy is looked up once per device with hipGetSymbolAddress,
// with that device current, and kept in noalloc_ptr[device].  Every call on
// one GPU shares that GPU's copy: calls that overlap on one GPU (concurrent
// streams or host threads) race.  Nothing in it has to start at zero: every
// block writes its own partial and the final pass reads exactly those.
#define PFHIP_NOALLOC_MAX_DEV 64
static __device__ __align__(256) double scratch[PFHIP_NOALLOC_SCRATCH_DOUBLES];
static double *noalloc_ptr[PFHIP_NOALLOC_MAX_DEV];
static pthread_mutex_t noalloc_mu = PTHREAD_MUTEX_INITIALIZER;

// the current device's copy of `scratch` (resolved on the first call per
// device); NULL with the error set on failure
static double *noalloc_scratch()
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= PFHIP_NOALLOC_MAX_DEV) {
    pfhip_set_error("no-allocation scratch: current device %d not usable", dev);
    return nullptr;
  }
  pthread_mutex_lock(&noalloc_mu);
  double *p = noalloc_ptr[dev];
  if (!p) {
    void *q = nullptr;
    hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q) {
      pfhip_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
    } else {
      p = (double *)q;
      noalloc_ptr[dev] = p;
    }
  }
  pthread_mutex_unlock(&noalloc_mu);
  return p;
}

extern "C" int pfhip_water_balance_storage_noalloc(pfhip_problem *pb, const double *p,
                                                   const double *s, double *subsurface,
                                                   double *surface, void *stream)
{
  if (!pb || !p || !s || !subsurface || !surface) {
    pfhip_set_error("water_balance_storage_noalloc: NULL argument");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  // layout: nb subsurface partials, nb surface partials, the two results
  if (2L * nb + 2 > PFHIP_NOALLOC_SCRATCH_DOUBLES) {
    pfhip_set_error("water_balance_storage_noalloc: %d reduction blocks need %ld doubles of "
                    "scratch, more than its %d", nb, 2L * nb + 2,
                    PFHIP_NOALLOC_SCRATCH_DOUBLES);
    return PFHIP_ERR_ARG;
  }
  double *sp = noalloc_scratch();
  if (!sp) return PFHIP_ERR_HIP;
  hipLaunchKernelGGL(k_storage, dim3(nb), dim3(256), 0, STREAM(stream), g, pb->d.dx,
                     pb->d.dy, pb->d.dz, p, s, pb->d.poro, pb->d.ss, sp, nb);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream), sp, nb, 2,
                     sp + 2 * nb);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(pb->h_red, sp + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost,
                        STREAM(stream)));
  HIPCHK(hipStreamSynchronize(STREAM(stream)));
  *subsurface = pb->h_red[0];
  *surface = pb->h_red[1];
  return PFHIP_OK;
}

// ---- block Gram-Schmidt kernels ------------------------------------------
#define MD_CHUNK 16

// the vectors of a multi-vector pass and a factor per vector: a pass reads
// sc[q] * v[q][i], the value a separate scaling pass (k_scale) would have
// stored, so a basis vector can be kept unnormalised with its 1/norm beside
// it (1.0 where a vector is used as stored: x * 1.0 is x)
struct VecPtrs {
  const double *v[MD_CHUNK];
  double sc[MD_CHUNK];
};
static VecPtrs vec_ptrs(const double *const *V, const double *scale, int c0, int mc,
                        const double *fill)
{
  VecPtrs P;
  for (int q = 0; q < MD_CHUNK; q++) {
    P.v[q] = (q < mc) ? V[c0 + q] : fill;
    P.sc[q] = (q < mc && scale) ? scale[c0 + q] : 1.0;
  }
  return P;
}

// The multi-vector kernels take the vector count M as a template argument:
// with a run-time count every load sat behind its own branch and waited for
// the one before it (one load in flight per wave); with M fixed the M loads
// of a cell are issued together.  md_launch picks the instantiation.
#define MD_CASE(n, ...) case n: __VA_ARGS__(n); break;
#define MD_SWITCH(m, L)                                                         \
  switch (m) {                                                                  \
    MD_CASE(1, L) MD_CASE(2, L) MD_CASE(3, L) MD_CASE(4, L) MD_CASE(5, L)       \
    MD_CASE(6, L) MD_CASE(7, L) MD_CASE(8, L) MD_CASE(9, L) MD_CASE(10, L)      \
    MD_CASE(11, L) MD_CASE(12, L) MD_CASE(13, L) MD_CASE(14, L) MD_CASE(15, L)  \
    MD_CASE(16, L)                                                              \
    default: break;                                                             \
  }

// h[j] = V[j].w for j < M (M <= MD_CHUNK), one pass over w
template <int M>
__global__ __launch_bounds__(256) void k_multi_dot(Geom g, VecPtrs V,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ partial, int nb)
{
  double acc[M];
#pragma unroll
  for (int q = 0; q < M; q++) acc[q] = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      double wv = w[ip];
      double v[M];
#pragma unroll
      for (int q = 0; q < M; q++) v[q] = V.sc[q] * V.v[q][ip];
#pragma unroll
      for (int q = 0; q < M; q++) acc[q] += v[q] * wv;
    }
  }
#pragma unroll
  for (int q = 0; q < M; q++) {
    double s = block_reduce<false>(acc[q]);
    if (threadIdx.x == 0) partial[q * nb + blockIdx.x] = s;
  }
}

static void launch_multi_dot(int mc, int nb, hipStream_t st, const Geom &g, const VecPtrs &P,
                             const double *w, double *partial)
{
#define L_MD(n) hipLaunchKernelGGL(k_multi_dot<n>, dim3(nb), dim3(256), 0, st, g, P, w, partial, nb)
  MD_SWITCH(mc, L_MD)
#undef L_MD
}

extern "C" int pfhip_vec_multi_dot(pfhip_problem *pb, const double *const *V,
                                   int m, const double *w, double *h, void *stream)
{
  if (m < 0 || m > PFHIP_MAX_MULTI) { pfhip_set_error("multi_dot m=%d", m); return PFHIP_ERR_ARG; }
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  for (int c0 = 0; c0 < m; c0 += MD_CHUNK) {
    int mc = (m - c0 < MD_CHUNK) ? m - c0 : MD_CHUNK;
    VecPtrs P = vec_ptrs(V, nullptr, c0, mc, w);
    launch_multi_dot(mc, nb, STREAM(stream), g, P, w, pb->red_partial + (long)c0 * nb);
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, m, pb->red_dev);
  LAUNCHCHK();
  int rc = fetch_results(pb, m, stream);
  if (rc) return rc;
  for (int q = 0; q < m; q++) h[q] = pb->h_red[q];
  return PFHIP_OK;
}

struct Coefs {
  double c[MD_CHUNK];
};

// w += sum_q c[q] V[q] (q < M); optional ||w||^2 partials
// w += sum_q C[q] V[q] (q < M); NORM: ||w||^2 partials; STEP (with NORM):
// also max |w| / (1 + |u|) partials (RED_SCSTEP's term with x = u, y = w) at
// partial[nb + block]; ZERO: w starts at 0 and is not read
template <int M, bool NORM, bool STEP = false, bool ZERO = false>
__global__ __launch_bounds__(256) void k_multi_axpy(Geom g, VecPtrs V, Coefs C,
                                                    double *__restrict__ w,
                                                    double *__restrict__ partial, int nb,
                                                    const double *__restrict__ u = nullptr)
{
  double acc = 0.0, mx = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      double v[M];
#pragma unroll
      for (int q = 0; q < M; q++) v[q] = V.sc[q] * V.v[q][ip];
      double s = ZERO ? 0.0 : w[ip];
#pragma unroll
      for (int q = 0; q < M; q++) s += C.c[q] * v[q];
      w[ip] = s;
      if (NORM) acc += s * s;
      if (STEP) mx = fmax(mx, fabs(s / (1.0 + fabs(u[ip]))));
    }
  }
  if (NORM) {
    double s = block_reduce<false>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
  if (STEP) {
    double m = block_reduce<true>(mx);
    if (threadIdx.x == 0) partial[nb + blockIdx.x] = m;
  }
}

static int multi_axpy_impl(pfhip_problem *pb, const double *const *V, int m,
                           const double *coef, double sign, double *w,
                           double *norm2_out, void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  for (int c0 = 0; c0 < m; c0 += MD_CHUNK) {
    int mc = (m - c0 < MD_CHUNK) ? m - c0 : MD_CHUNK;
    VecPtrs P = vec_ptrs(V, nullptr, c0, mc, w);
    Coefs C;
    for (int q = 0; q < MD_CHUNK; q++) C.c[q] = (q < mc) ? sign * coef[c0 + q] : 0.0;
    bool last = (c0 + MD_CHUNK >= m);
    hipStream_t st = STREAM(stream);
    if (last && norm2_out) {
#define L_MA(n) hipLaunchKernelGGL((k_multi_axpy<n, true>), dim3(nb), dim3(256), 0, st, g, P, C, w, pb->red_partial, nb)
      MD_SWITCH(mc, L_MA)
#undef L_MA
    } else {
#define L_MA(n) hipLaunchKernelGGL((k_multi_axpy<n, false>), dim3(nb), dim3(256), 0, st, g, P, C, w, pb->red_partial, nb)
      MD_SWITCH(mc, L_MA)
#undef L_MA
    }
    LAUNCHCHK();
  }
  if (norm2_out) {
    if (m == 0) return pfhip_vec_dot(pb, w, w, norm2_out, stream);
    hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                       pb->red_partial, nb, 1, pb->red_dev);
    LAUNCHCHK();
    int rc = fetch_results(pb, 1, stream);
    if (rc) return rc;
    *norm2_out = pb->h_red[0];
  }
  return PFHIP_OK;
}

// w -= sum_q h[q] V[q] (q < M) with h read from device memory (the
// coefficients the previous multi-dot left there); optional ||w||^2 partials
template <int M, bool NORM>
__global__ __launch_bounds__(256) void k_multi_axpy_d(Geom g, VecPtrs V,
                                                      const double *__restrict__ h,
                                                      double *__restrict__ w,
                                                      double *__restrict__ partial, int nb)
{
  double c[M];
#pragma unroll
  for (int q = 0; q < M; q++) c[q] = -1.0 * h[q];
  double acc = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      double v[M];
#pragma unroll
      for (int q = 0; q < M; q++) v[q] = V.sc[q] * V.v[q][ip];
      double s = w[ip];
#pragma unroll
      for (int q = 0; q < M; q++) s += c[q] * v[q];
      w[ip] = s;
      if (NORM) acc += s * s;
    }
  }
  if (NORM) {
    double s = block_reduce<false>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// CGS2's middle passes in one: w -= V h (h on the device), then the dot
// products V[q].w of the updated w, per cell in the order of k_multi_axpy_d
// followed by k_multi_dot -- the same products and sums, one read of V
template <int M>
__global__ __launch_bounds__(256) void k_multi_axpy_dot(Geom g, VecPtrs V,
                                                        const double *__restrict__ h,
                                                        double *__restrict__ w,
                                                        double *__restrict__ partial, int nb)
{
  double c[M], acc[M];
#pragma unroll
  for (int q = 0; q < M; q++) { c[q] = -1.0 * h[q]; acc[q] = 0.0; }
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      double v[M];
#pragma unroll
      for (int q = 0; q < M; q++) v[q] = V.sc[q] * V.v[q][ip];
      double s = w[ip];
#pragma unroll
      for (int q = 0; q < M; q++) s += c[q] * v[q];
      w[ip] = s;
#pragma unroll
      for (int q = 0; q < M; q++) acc[q] += v[q] * s;
    }
  }
#pragma unroll
  for (int q = 0; q < M; q++) {
    double t = block_reduce<false>(acc[q]);
    if (threadIdx.x == 0) partial[q * nb + blockIdx.x] = t;
  }
}

// one device pass of h = V^T w (m values into dst)
static void multi_dot_dev(pfhip_problem *pb, const Geom &g, int nb, const double *const *V,
                          const double *sc, int m, const double *w, double *dst,
                          hipStream_t st)
{
  for (int c0 = 0; c0 < m; c0 += MD_CHUNK) {
    int mc = (m - c0 < MD_CHUNK) ? m - c0 : MD_CHUNK;
    VecPtrs P = vec_ptrs(V, sc, c0, mc, w);
    launch_multi_dot(mc, nb, st, g, P, w, pb->red_partial + (long)c0 * nb);
  }
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, st, pb->red_partial, nb, m,
                     dst);
}

// w -= V h with h on the device; the last chunk optionally adds ||w||^2
// partials, reduced into *n2dst
static void multi_axpy_dev(pfhip_problem *pb, const Geom &g, int nb, const double *const *V,
                           const double *sc, int m, const double *h, double *w,
                           double *n2dst, hipStream_t st)
{
  for (int c0 = 0; c0 < m; c0 += MD_CHUNK) {
    int mc = (m - c0 < MD_CHUNK) ? m - c0 : MD_CHUNK;
    VecPtrs P = vec_ptrs(V, sc, c0, mc, w);
    bool last = (c0 + MD_CHUNK >= m);
    if (last && n2dst) {
#define L_MAD(n) hipLaunchKernelGGL((k_multi_axpy_d<n, true>), dim3(nb), dim3(256), 0, st, g, P, h + c0, w, pb->red_partial, nb)
      MD_SWITCH(mc, L_MAD)
#undef L_MAD
    } else {
#define L_MAD(n) hipLaunchKernelGGL((k_multi_axpy_d<n, false>), dim3(nb), dim3(256), 0, st, g, P, h + c0, w, pb->red_partial, nb)
      MD_SWITCH(mc, L_MAD)
#undef L_MAD
    }
  }
  if (n2dst)
    hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, st, pb->red_partial, nb, 1,
                       n2dst);
}

// CGS2 against V[0..k-1] with one host synchronisation: h1 = V^T w,
// w -= V h1, h2 = V^T w, w -= V h2 and ||w||^2, the coefficients passing
// from each reduction to the next update on the device.  The arithmetic is
// that of the separate multi-dot / multi-axpy calls.  sc (or null): basis
// vector q is read as sc[q] * V[q], the vector a k_scale pass by sc[q]
// would have stored.
int pfhip_vec_cgs2_scaled(pfhip_problem *pb, const double *const *V, const double *sc, int k,
                          double *w, double *h1, double *h2, double *norm2, void *stream)
{
  if (k < 1 || k > PFHIP_SLOT_H2) {
    pfhip_set_error("cgs2: k=%d outside 1..%d", k, PFHIP_SLOT_H2);
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipStream_t st = STREAM(stream);
  double *d1 = pb->red_dev, *d2 = pb->red_dev + PFHIP_SLOT_H2;
  double *dn = pb->red_dev + PFHIP_SLOT_N2;
  multi_dot_dev(pb, g, nb, V, sc, k, w, d1, st);
  if (k <= 8) {
    // w -= V h1 and h2 = V^T w in one pass over V
    VecPtrs P = vec_ptrs(V, sc, 0, k, w);
#define L_MADOT(n) hipLaunchKernelGGL(k_multi_axpy_dot<n>, dim3(nb), dim3(256), 0, st, g, P, d1, w, pb->red_partial, nb)
    switch (k) {
      MD_CASE(1, L_MADOT) MD_CASE(2, L_MADOT) MD_CASE(3, L_MADOT) MD_CASE(4, L_MADOT)
      MD_CASE(5, L_MADOT) MD_CASE(6, L_MADOT) MD_CASE(7, L_MADOT) MD_CASE(8, L_MADOT)
      default: break;
    }
#undef L_MADOT
    hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, st, pb->red_partial, nb,
                       k, d2);
  } else {
    multi_axpy_dev(pb, g, nb, V, sc, k, d1, w, nullptr, st);
    multi_dot_dev(pb, g, nb, V, sc, k, w, d2, st);
  }
  multi_axpy_dev(pb, g, nb, V, sc, k, d2, w, dn, st);
  LAUNCHCHK();
  int rc = fetch_results(pb, PFHIP_MAX_MULTI, stream);
  if (rc) return rc;
  for (int q = 0; q < k; q++) { h1[q] = pb->h_red[q]; h2[q] = pb->h_red[PFHIP_SLOT_H2 + q]; }
  *norm2 = pb->h_red[PFHIP_SLOT_N2];
  return PFHIP_OK;
}

int pfhip_vec_cgs2(pfhip_problem *pb, const double *const *V, int k, double *w, double *h1,
                   double *h2, double *norm2, void *stream)
{
  return pfhip_vec_cgs2_scaled(pb, V, nullptr, k, w, h1, h2, norm2, stream);
}

extern "C" int pfhip_vec_multi_axpy(pfhip_problem *pb, const double *const *V,
                                    int m, const double *h, double *w,
                                    double *norm2_out, void *stream)
{
  return multi_axpy_impl(pb, V, m, h, -1.0, w, norm2_out, stream);
}

extern "C" int pfhip_vec_combine(pfhip_problem *pb, const double *const *V, int m,
                                 const double *y, double *x, void *stream)
{
  return multi_axpy_impl(pb, V, m, y, 1.0, x, nullptr, stream);
}

// x = sum_k y[k] V[k] (zero != 0: x's contents are not read) or x +=, and
// the line search's two terms of the result from the same pass: x.x (the
// products and order of pfhip_vec_dot(x, x)) and max |x| / (1 + |u|) (those
// of pfhip_vec_scsteplength(u, x)); one synchronisation
extern "C" int pfhip_vec_combine_step(pfhip_problem *pb, const double *const *V, int m,
                                      const double *y, double *x, const double *u, int zero,
                                      double *xx, double *steplength, void *stream)
{
  if (m < 1) {
    pfhip_set_error("vec_combine_step: no vectors");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipStream_t st = STREAM(stream);
  for (int c0 = 0; c0 < m; c0 += MD_CHUNK) {
    int mc = (m - c0 < MD_CHUNK) ? m - c0 : MD_CHUNK;
    VecPtrs P = vec_ptrs(V, nullptr, c0, mc, x);
    Coefs C;
    for (int q = 0; q < MD_CHUNK; q++) C.c[q] = (q < mc) ? y[c0 + q] : 0.0;
    const bool last = (c0 + MD_CHUNK >= m), z = zero && c0 == 0;
    if (last && z) {
#define L_MA(n) hipLaunchKernelGGL((k_multi_axpy<n, true, true, true>), dim3(nb), dim3(256), 0, st, g, P, C, x, pb->red_partial, nb, u)
      MD_SWITCH(mc, L_MA)
#undef L_MA
    } else if (last) {
#define L_MA(n) hipLaunchKernelGGL((k_multi_axpy<n, true, true, false>), dim3(nb), dim3(256), 0, st, g, P, C, x, pb->red_partial, nb, u)
      MD_SWITCH(mc, L_MA)
#undef L_MA
    } else if (z) {
#define L_MA(n) hipLaunchKernelGGL((k_multi_axpy<n, false, false, true>), dim3(nb), dim3(256), 0, st, g, P, C, x, pb->red_partial, nb, nullptr)
      MD_SWITCH(mc, L_MA)
#undef L_MA
    } else {
#define L_MA(n) hipLaunchKernelGGL((k_multi_axpy<n, false>), dim3(nb), dim3(256), 0, st, g, P, C, x, pb->red_partial, nb, nullptr)
      MD_SWITCH(mc, L_MA)
#undef L_MA
    }
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, st, pb->red_partial, nb, 1,
                     pb->red_dev);
  hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(256), 0, st, pb->red_partial + nb,
                     nb, 1, pb->red_dev + 1);
  LAUNCHCHK();
  int rc = fetch_results(pb, 2, stream);
  if (rc) return rc;
  *xx = pb->h_red[0];
  *steplength = pb->h_red[1];
  return PFHIP_OK;
}

// MGS step: w -= hprev*vprev (if vprev) and dot = vnext.w (if vnext)
__global__ __launch_bounds__(256) void k_axpy_dot(Geom g, double hprev,
                                                  const double *__restrict__ vprev,
                                                  double *__restrict__ w,
                                                  const double *__restrict__ vnext,
                                                  double *__restrict__ partial)
{
  double acc = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      int ip = base + i;
      double s = w[ip];
      if (vprev) { s = 1.0 * s + (-hprev) * vprev[ip]; w[ip] = s; }
      if (vnext) acc += vnext[ip] * s;
    }
  }
  double s = block_reduce<false>(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

extern "C" int pfhip_vec_axpy_dot(pfhip_problem *pb, double hprev,
                                  const double *vprev, double *w,
                                  const double *vnext, double *dot, void *stream)
{
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_axpy_dot, dim3(nb), dim3(256), 0, STREAM(stream), g, hprev,
                     vprev, w, vnext, pb->red_partial);
  LAUNCHCHK();
  if (!vnext) return PFHIP_OK;
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 1, pb->red_dev);
  LAUNCHCHK();
  int rc = fetch_results(pb, 1, stream);
  if (rc) return rc;
  *dot = pb->h_red[0];
  return PFHIP_OK;
}

// ---- elementwise ----------------------------------------------------------
__global__ __launch_bounds__(256) void k_linear_sum(Geom g, double a,
                                                    const double *__restrict__ x,
                                                    double b,
                                                    const double *__restrict__ y,
                                                    double *__restrict__ z)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  z[ip] = a * x[ip] + b * y[ip];
}

__global__ __launch_bounds__(256) void k_scale(Geom g, double c,
                                               const double *__restrict__ x,
                                               double *__restrict__ z)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  z[ip] = c * x[ip];
}

__global__ __launch_bounds__(256) void k_const(Geom g, double c, double *__restrict__ z)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  z[(i + 1) + (j + 1) * g.sy + (k + 1) * g.sz] = c;
}

extern "C" int pfhip_vec_linear_sum(pfhip_problem *pb, double a, const double *x,
                                    double b, const double *y, double *z, void *stream)
{
  Geom g = geom_of(pb);
  hipLaunchKernelGGL(k_linear_sum, grid3(g), BLOCK3, 0, STREAM(stream), g, a, x, b, y, z);
  LAUNCHCHK();
  return PFHIP_OK;
}

extern "C" int pfhip_vec_scale(pfhip_problem *pb, double c, const double *x,
                               double *z, void *stream)
{
  Geom g = geom_of(pb);
  hipLaunchKernelGGL(k_scale, grid3(g), BLOCK3, 0, STREAM(stream), g, c, x, z);
  LAUNCHCHK();
  return PFHIP_OK;
}

extern "C" int pfhip_vec_const(pfhip_problem *pb, double c, double *z, void *stream)
{
  Geom g = geom_of(pb);
  hipLaunchKernelGGL(k_const, grid3(g), BLOCK3, 0, STREAM(stream), g, c, z);
  LAUNCHCHK();
  return PFHIP_OK;
}

// The other N_Vector operations of PL/n_vector.h (PL/vector_utilities.c),
// elementwise over the interior cells with the reference's operation per
// cell; ghost entries are left as they are.
enum EwOp { EW_PROD = 0, EW_DIV = 1, EW_ABS = 2, EW_INV = 3, EW_ADDC = 4, EW_CMP = 5 };

template <int OP>
__global__ __launch_bounds__(256) void k_vec_ew(Geom g, double c, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                double *__restrict__ z)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  double v = x[ip];
  if (OP == EW_PROD) z[ip] = v * y[ip];
  else if (OP == EW_DIV) z[ip] = v / y[ip];
  else if (OP == EW_ABS) z[ip] = fabs(v);
  else if (OP == EW_INV) z[ip] = 1.0 / v;
  else if (OP == EW_ADDC) z[ip] = v + c;
  else z[ip] = (fabs(v) >= c) ? 1.0 : 0.0;
}

template <int OP>
static int vec_ew(pfhip_problem *pb, double c, const double *x, const double *y, double *z,
                  void *stream)
{
  if (!pb || !x || !z || ((OP == EW_PROD || OP == EW_DIV) && !y)) {
    pfhip_set_error("N_Vector operation: NULL argument");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  hipLaunchKernelGGL(k_vec_ew<OP>, grid3(g), BLOCK3, 0, STREAM(stream), g, c, x, y, z);
  LAUNCHCHK();
  return PFHIP_OK;
}

extern "C" int pfhip_vec_prod(pfhip_problem *pb, const double *x, const double *y, double *z,
                              void *stream)
{
  return vec_ew<EW_PROD>(pb, 0.0, x, y, z, stream);
}
extern "C" int pfhip_vec_div(pfhip_problem *pb, const double *x, const double *y, double *z,
                             void *stream)
{
  return vec_ew<EW_DIV>(pb, 0.0, x, y, z, stream);
}
extern "C" int pfhip_vec_abs(pfhip_problem *pb, const double *x, double *z, void *stream)
{
  return vec_ew<EW_ABS>(pb, 0.0, x, nullptr, z, stream);
}
extern "C" int pfhip_vec_inv(pfhip_problem *pb, const double *x, double *z, void *stream)
{
  return vec_ew<EW_INV>(pb, 0.0, x, nullptr, z, stream);
}
extern "C" int pfhip_vec_add_const(pfhip_problem *pb, const double *x, double b, double *z,
                                   void *stream)
{
  return vec_ew<EW_ADDC>(pb, b, x, nullptr, z, stream);
}
extern "C" int pfhip_vec_compare(pfhip_problem *pb, double c, const double *x, double *z,
                                 void *stream)
{
  return vec_ew<EW_CMP>(pb, c, x, nullptr, z, stream);
}

// InvTest: z = 1/x where x != 0 (z keeps its value where x == 0), and a
// count of the zeros; row-owning blocks as the reductions
__global__ __launch_bounds__(256) void k_inv_test(Geom g, const double *__restrict__ x,
                                                  double *__restrict__ z,
                                                  double *__restrict__ partial)
{
  double zeros = 0.0;
  int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    int j = r % g.ny, k = r / g.ny;
    int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      double v = x[base + i];
      if (v == 0.0) zeros += 1.0;
      else z[base + i] = 1.0 / v;
    }
  }
  zeros = block_reduce<false>(zeros);
  if (threadIdx.x == 0) partial[blockIdx.x] = zeros;
}

// --------------------------------------------------------------------------
// van Genuchten closures: S and dS/dp (PL/problem_saturation.c:265-310),
// kr and dkr/dp (PL/problem_phase_rel_perm.c:767-792, 895-912).
//
// The reference spends 3-8 pow() calls per cell.  Here every power of
// (alpha h) and of (1 + (alpha h)^n) is derived from one shared chain,
//   L = log(alpha h), t = (alpha h)^n = exp(n L),
//   sa = (1+t)^(-m/2) = exp(-m/2 log1p(t)), a = (1+t)^-m = sa^2,
// so S and kr together cost two logs and two exps (pf_math.h); the
// derivative terms reuse t, a and sa.  h = |p| / (rho g) takes 1/(rho g)
// from the caller.  Values agree with the pow form to a few ulp (the
// 1 - (alpha h)^(n-1) a cancellation of dry cells is common to both forms).
// --------------------------------------------------------------------------
struct VGTerms {
  double ah, t, a, sa, ahnm1, m;
};

__device__ __forceinline__ VGTerms vg_terms(double p, double irg, double alpha, double n)
{
  VGTerms v;
  v.m = 1.0 - pf_rcp(n);
  double head = fabs(p) * irg;
  v.ah = alpha * head;
  double L = pf_log(v.ah);
  v.t = pf_exp(n * L);                            // (alpha h)^n
  v.sa = pf_exp(-0.5 * v.m * pf_log1p(v.t));      // (1 + (alpha h)^n)^(-m/2)
  v.a = v.sa * v.sa;                              // (1 + (alpha h)^n)^-m
  v.ahnm1 = pf_div(v.t, v.ah);                    // (alpha h)^(n-1)
  return v;
}

__device__ __forceinline__ double vg_sat(double p, double irg, double alpha, double n,
                                         double sres, double ssat)
{
  double s_dif = ssat - sres;
  if (p >= 0.0) return s_dif + sres;
  VGTerms v = vg_terms(p, irg, alpha, n);
  return s_dif * v.a + sres;
}

__device__ __forceinline__ double vg_dsat(double p, double irg, double alpha, double n,
                                          double sres, double ssat)
{
  double s_dif = ssat - sres;
  if (p >= 0.0) return 0.0;
  VGTerms v = vg_terms(p, irg, alpha, n);
  return (v.m * n * alpha * v.ahnm1) * s_dif * pf_div(v.a, 1.0 + v.t);
}

// kr lookup table of PL/problem_phase_rel_perm.c:116-354 (region-uniform
// alpha, n): 7 arrays of n+1 values x | a | d | a_der | d_der | slope |
// slope_der on the device; t == nullptr selects the closed form
struct KrTab {
  const double *__restrict__ t;
  int n, method;                     // method 0 monotone cubic Hermite, 1 linear
  double interval, maxhead;          // node spacing, |MinPressureHead|
};

// kr (der = false) or the derivative formula's value (der = true) at head
// h >= 0: VanGLookupSpline (:265-354) / the linear lookup (:718-760)
__device__ __forceinline__ double kr_table(const KrTab &tb, double head, bool der)
{
  if (head >= tb.maxhead) return 0.0;
  const int N1 = tb.n + 1;
  int pt = (int)floor(head / tb.interval);
  const double *x = tb.t;
  const double *a = tb.t + N1 * (der ? 3 : 1);
  if (tb.method == 1) {
    const double *sl = tb.t + N1 * (der ? 6 : 5);
    return a[pt] + sl[pt] * (head - x[pt]);
  }
  if (pt > tb.n + 1) pt = tb.n;
  const double *d = tb.t + N1 * (der ? 4 : 2);
  const double x0 = x[pt], h = x[pt + 1] - x0;
  const double t = (head - x0) / h, t2 = t * t, t3 = t2 * t;
  return (2.0 * t3 - 3.0 * t2 + 1.0) * a[pt] + (t3 - 2.0 * t2 + t) * h * d[pt] +
         (-2.0 * t3 + 3.0 * t2) * a[pt + 1] + (t3 - t2) * h * d[pt + 1];
}

__device__ __forceinline__ double vg_kr_t(const VGTerms &v)
{
  double coeff = 1.0 - v.ahnm1 * v.a;
  return (coeff * coeff) * v.sa;
}

__device__ __forceinline__ double vg_dkr_t(const VGTerms &v, double alpha, double n)
{
  double m = v.m, ropahn = pf_rcp(1.0 + v.t);
  double coeff = 1.0 - v.ahnm1 * v.a;
  return 2.0 * (coeff * v.sa) *
             ((n - 1) * pf_div(v.ahnm1, v.ah) * alpha * v.a -
              v.ahnm1 * m * (v.a * ropahn) * n * alpha * v.ahnm1) +
         (coeff * coeff) * (m / 2) * (v.sa * ropahn) * n * alpha * v.ahnm1;
}

template <bool TAB = true>
__device__ __forceinline__ double vg_kr(double p, double irg, double alpha, double n,
                                        const KrTab &tb)
{
  if (p >= 0.0) return 1.0;
  if (TAB && tb.t) return kr_table(tb, fabs(p) * irg, false);
  return vg_kr_t(vg_terms(p, irg, alpha, n));
}

template <bool TAB = true>
__device__ __forceinline__ double vg_dkr(double p, double irg, double alpha, double n,
                                         const KrTab &tb)
{
  if (p >= 0.0) return 0.0;
  if (TAB && tb.t) return kr_table(tb, fabs(p) * irg, true);
  return vg_dkr_t(vg_terms(p, irg, alpha, n), alpha, n);
}

// S and kr of one cell from one chain; TAB = false compiles the closed form
// only (kernels specialised for problems without a table)
template <bool TAB = true>
__device__ __forceinline__ void vg_s_kr(double p, double irg, double alpha, double n,
                                        double sres, double ssat, double &S, double &kr,
                                        const KrTab &tb)
{
  double s_dif = ssat - sres;
  if (p >= 0.0) { S = s_dif + sres; kr = 1.0; return; }
  VGTerms v = vg_terms(p, irg, alpha, n);
  S = s_dif * v.a + sres;
  kr = (TAB && tb.t) ? kr_table(tb, fabs(p) * irg, false) : vg_kr_t(v);
}

struct Phys {
  double rho, mu, g;
  const double *__restrict__ alpha;
  const double *__restrict__ nvg;
  const double *__restrict__ sres;   // NULL: uniform sres_v
  const double *__restrict__ ssat;   // NULL: uniform ssat_v
  double sres_v, ssat_v;
  double irg;                        // 1 / (rho g)
  KrTab tb;                          // kr lookup table or tb.t == nullptr
  __device__ __forceinline__ double Sres(int ip) const { return sres ? sres[ip] : sres_v; }
  __device__ __forceinline__ double Ssat(int ip) const { return ssat ? ssat[ip] : ssat_v; }
};

static inline Phys phys_of(const pfhip_problem *pb)
{
  Phys f;
  f.rho = pb->d.rho; f.mu = pb->d.mu; f.g = pb->d.gravity;
  f.alpha = pb->d.alpha; f.nvg = pb->d.nvg; f.sres = pb->d.sres; f.ssat = pb->d.ssat;
  f.sres_v = pb->d.sres_value; f.ssat_v = pb->d.ssat_value;
  f.irg = 1.0 / (f.rho * f.g);
  f.tb.t = pb->krtab;
  f.tb.n = pb->d.vg_table_points;
  f.tb.method = pb->d.vg_table_method;
  f.tb.interval = fabs(pb->d.vg_table_min_head / (double)(pb->d.vg_table_points - 1));
  f.tb.maxhead = fabs(pb->d.vg_table_min_head);
  return f;
}

// out = S (fcn 0) / dS (1) / kr (2) / dkr (3) / kr*rho (4), or both
// kr*rho and dkr*rho into out, out2 (5); interior cells
__global__ __launch_bounds__(256) void k_closure(Geom g, Phys f, int what,
                                                 const double *__restrict__ p,
                                                 double *__restrict__ out,
                                                 double *__restrict__ out2)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  double pv = p[ip], a = f.alpha[ip], n = f.nvg[ip];
  switch (what) {
    case 0: out[ip] = vg_sat(pv, f.irg, a, n, f.Sres(ip), f.Ssat(ip)); break;
    case 1: out[ip] = vg_dsat(pv, f.irg, a, n, f.Sres(ip), f.Ssat(ip)); break;
    case 2: out[ip] = vg_kr(pv, f.irg, a, n, f.tb); break;
    case 3: out[ip] = vg_dkr(pv, f.irg, a, n, f.tb); break;
    case 4: out[ip] = vg_kr(pv, f.irg, a, n, f.tb) * f.rho; break;
    default: {
      double kr = vg_kr(pv, f.irg, a, n, f.tb);
      out[ip] = kr * f.rho;
      out2[ip] = vg_dkr(pv, f.irg, a, n, f.tb) * f.rho + kr * 0.0;
    }
  }
}

static int launch_closure(pfhip_problem *pb, int what, const double *p, double *out,
                          double *out2, void *stream)
{
  Geom g = geom_of(pb);
  PROF_BEGIN(PFHIP_PROF_CLOSURE, stream);
  hipLaunchKernelGGL(k_closure, grid3(g), BLOCK3, 0, STREAM(stream), g, phys_of(pb),
                     what, p, out, out2);
  PROF_END(PFHIP_PROF_CLOSURE, stream);
  LAUNCHCHK();
  return PFHIP_OK;
}

extern "C" int pfhip_saturation(pfhip_problem *pb, const double *p, double *out,
                                int fcn, void *stream)
{
  return launch_closure(pb, fcn == PFHIP_CALCDER ? 1 : 0, p, out, nullptr, stream);
}

extern "C" int pfhip_phase_rel_perm(pfhip_problem *pb, const double *p, double *out,
                                    int fcn, void *stream)
{
  return launch_closure(pb, fcn == PFHIP_CALCDER ? 3 : 2, p, out, nullptr, stream);
}

// --------------------------------------------------------------------------
// kinematic overland flux (PL/overlandflow_eval_Kin.c:120-460), one thread
// per surface column of a box whose whole top is the overland patch
// --------------------------------------------------------------------------
struct Surf {
  const double *__restrict__ sx;
  const double *__restrict__ sy;
  const double *__restrict__ mann;
  double eps;
};

__device__ __forceinline__ double rpowerr(double b, double e)
{
  return (b <= 0.0) ? 0.0 : pow(b, e);
}

__global__ __launch_bounds__(256) void k_overland_q(Geom g, Surf s,
                                                    const double *__restrict__ p,
                                                    double *__restrict__ qx,
                                                    double *__restrict__ qy)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  int io = (i + 1) + (j + 1) * g.sy;
  int kt = g.nz - 1;
  int ip = io + (kt + 1) * g.sz;
  double Sf_x = s.sx[io], Sf_y = s.sy[io];
  double Sf_mag = rpowerr(Sf_x * Sf_x + Sf_y * Sf_y, 0.5);
  if (Sf_mag < s.eps) Sf_mag = s.eps;
  double pc = p[ip];
  double pe = (i + 1 < g.nx) ? p[ip + 1] : 0.0;
  double pn = (j + 1 < g.ny) ? p[ip + g.sy] : 0.0;
  double Px = ((-Sf_x - 0.0) >= 0) ? fmax(pc, 0.0) : fmax(pe, 0.0);
  double Py = ((-Sf_y - 0.0) >= 0) ? fmax(pc, 0.0) : fmax(pn, 0.0);
  double den = rpowerr(fabs(Sf_mag), 0.5) * s.mann[io];
  qx[io] = -(Sf_x / den) * rpowerr(Px, (5.0 / 3.0));
  qy[io] = -(Sf_y / den) * rpowerr(Py, (5.0 / 3.0));
  if (i == 0)
    qx[io - 1] = (Sf_x > 0.0) ? -(Sf_x / den) * rpowerr(fmax(pc, 0.0), (5.0 / 3.0)) : 0.0;
  if (j == 0)
    qy[io - g.sy] = (Sf_y > 0.0) ? -(Sf_y / den) * rpowerr(fmax(pc, 0.0), (5.0 / 3.0)) : 0.0;
}

__global__ __launch_bounds__(256) void k_overland_der(Geom g, Surf s,
                                                      const double *__restrict__ p,
                                                      double *__restrict__ ke,
                                                      double *__restrict__ kw,
                                                      double *__restrict__ kn,
                                                      double *__restrict__ ks)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  int io = (i + 1) + (j + 1) * g.sy;
  int kt = g.nz - 1;
  int ip = io + (kt + 1) * g.sz;
  double Sf_x = s.sx[io], Sf_y = s.sy[io];
  double Sf_mag = rpowerr(Sf_x * Sf_x + Sf_y * Sf_y, 0.5);
  if (Sf_mag < s.eps) Sf_mag = s.eps;
  double pc = p[ip];
  double pe = (i + 1 < g.nx) ? p[ip + 1] : 0.0;
  double pn = (j + 1 < g.ny) ? p[ip + g.sy] : 0.0;
  double Px = ((-Sf_x - 0.0) >= 0) ? fmax(pc, 0.0) : fmax(pe, 0.0);
  double Py = ((-Sf_y - 0.0) >= 0) ? fmax(pc, 0.0) : fmax(pn, 0.0);
  double den = rpowerr(fabs(Sf_mag), 0.5) * s.mann[io];
  double qxt = -(5.0 / 3.0) * (Sf_x / den) * rpowerr(Px, (2.0 / 3.0));
  double qyt = -(5.0 / 3.0) * (Sf_y / den) * rpowerr(Py, (2.0 / 3.0));
  ke[io] = fmax(qxt, 0.0);
  kw[io + 1] = -fmax(-qxt, 0.0);
  kn[io] = fmax(qyt, 0.0);
  ks[io + g.sy] = -fmax(-qyt, 0.0);
  if (i == 0 && Sf_x > 0.0) {
    double q = -(5.0 / 3.0) * (Sf_x / den) * rpowerr(fmax(pc, 0.0), (2.0 / 3.0));
    kw[io] = -fmax(-q, 0.0);
    ke[io - 1] = fmax(q, 0.0);
  }
  if (j == 0 && Sf_y > 0.0) {
    double q = -(5.0 / 3.0) * (Sf_y / den) * rpowerr(fmax(pc, 0.0), (2.0 / 3.0));
    ks[io] = -fmax(-q, 0.0);
    kn[io - g.sy] = fmax(q, 0.0);
  }
}

static inline Surf surf_of(const pfhip_problem *pb)
{
  Surf s;
  s.sx = pb->d.slope_x; s.sy = pb->d.slope_y; s.mann = pb->d.mann; s.eps = pb->d.ov_eps;
  return s;
}

static inline dim3 grid2(const Geom &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, 1); }

extern "C" int pfhip_overland_flow_eval_kin(pfhip_problem *pb, const double *p,
                                            double *o0, double *o1, double *o2,
                                            double *o3, int fcn, void *stream)
{
  if (!pb->d.slope_x) { pfhip_set_error("problem has no surface fields"); return PFHIP_ERR_ARG; }
  Geom g = geom_of(pb);
  if (fcn == PFHIP_CALCFCN) {
    hipLaunchKernelGGL(k_overland_q, grid2(g), BLOCK3, 0, STREAM(stream), g,
                       surf_of(pb), p, o0, o1);
  } else {
    for (double *a : {o0, o1, o2, o3}) {
      int rc = pfhip_zero(a, pb->n2, stream);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_overland_der, grid2(g), BLOCK3, 0, STREAM(stream), g,
                       surf_of(pb), p, o0, o1, o2, o3);
  }
  LAUNCHCHK();
  return PFHIP_OK;
}

// --------------------------------------------------------------------------
// residual F(p) -- NlFunctionEval, PL/nl_function_eval.c:167-2323, gather
// form: each cell sums the flux through its six faces.  A face flux is
// computed from both sides with identical operands and operation order, so
// both cells see the bit-identical value the reference scatters with
// PlusEquals; no atomics.  Ghost permeability is zero in the reference
// (PL/vector.c:321), so a domain face contributes only its BC "u_new" term
// (:975-2260), which is what the boundary branch computes.
// --------------------------------------------------------------------------
// Tile order of the residual and Jacobian launches.  Workgroups are dealt
// to the 8 XCDs round-robin, so in launch order the x-neighbouring tiles,
// which share the cache lines at their x edges (a tile row of 64 FP64
// values straddles five 128-byte lines in the ghosted layout), would land
// on different XCDs with separate L2s.  Here each run of 8 R gx launch
// slots gives every XCD R whole tile rows (all gx tiles of a row): the
// shared lines are then fetched once per XCD.  Tails that do not fill a
// run keep the launch order.  Measured on 16.8M cells: F -2.5 %, the
// line-search form -2.7 %, the Jacobian -4 %, but the J.v form +2.5 %, which
// therefore keeps the launch order (REMAP = false).
#define TILE_ROWS_PER_XCD 4
template <bool REMAP = true>
__device__ __forceinline__ void tile_of_block(int &bx, int &by, int &bz)
{
  if (!REMAP) { bx = blockIdx.x; by = blockIdx.y; bz = blockIdx.z; return; }
  const int gx = gridDim.x, gy = gridDim.y;
  const int N = gx * gy * gridDim.z, G = 8 * TILE_ROWS_PER_XCD * gx;
  const int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int q = b / G, r = b - q * G;
  const int t = ((q + 1) * G <= N) ? q * G + (r & 7) * TILE_ROWS_PER_XCD * gx + (r >> 3) : b;
  bx = t % gx;
  const int row = t / gx;
  by = row % gy;
  bz = row / gy;
}

struct ResArgs {
  Geom g;
  Phys f;
  double dx, dy, dz, dt;
  double idx, idy, idz, imu;         // reciprocals of dx, dy, dz, mu
  const double *__restrict__ p;
  // k_residual_pf forms: MODE 1 (difference-quotient J.v) evaluates at
  // p + psig pv and writes fsc F - fsc fsub; MODE 2 (line-search step)
  // evaluates at p + psig pv, writes F and the pressure itself to pout
  const double *__restrict__ pv;
  const double *__restrict__ fsub;
  double *__restrict__ pout;
  double psig, fsc;
  // MODE 1 with dq set: sigma and 1/sigma of KINSpgmrAtimesDQ from the
  // device sums dq = (u.v, v.v, |v|_1) and sqrt(relfunc) sqrf, in the kernel
  const double *__restrict__ dq;
  double sqrf;
  const double *__restrict__ p_old;
  const double *__restrict__ s_old;
  const double *__restrict__ kx;
  const double *__restrict__ ky;
  const double *__restrict__ kz;
  const double *__restrict__ poro;
  const double *__restrict__ ss;
  const double *__restrict__ src;
  const double *__restrict__ krho;
  const double *__restrict__ qx;
  const double *__restrict__ qy;
  const double *bcv[6];
  int bct[6];
  double *__restrict__ F;
  // MODE 2 with stat set: per-block KINStop maxima (max |F|, max scaled
  // step) into stat[blk] and stat[stat_nb + blk]; the tiled kernel's blocks
  // first (interior cells), then the boundary kernel's from stat_main on
  double *stat;
  long stat_nb, stat_main;
};

// max over a workgroup of NT threads of two values; valid in thread 0
template <int NT>
__device__ __forceinline__ void block_max2(double &a, double &b)
{
  __shared__ double sh[2][NT / 64];
  a = wave_max(a);
  b = wave_max(b);
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  if ((t & 63) == 0) { sh[0][t >> 6] = a; sh[1][t >> 6] = b; }
  __syncthreads();
  if (t == 0)
    for (int w = 1; w < NT / 64; w++) { a = fmax(a, sh[0][w]); b = fmax(b, sh[1][w]); }
  __syncthreads();
}

__device__ __forceinline__ double hmean(double a, double b)
{
  return (a + b) ? (2.0 * a * b) / (a + b) : 0;
}
__device__ __forceinline__ double hmean_dz1(double a, double b)
{
  // HarmonicMeanDZ(a, b, 1, 1), PL/solver.h:45
  return ((1.0 * b) + (a * 1.0)) ? (((1.0 + 1.0) * a * b) / ((b * 1.0) + (a * 1.0))) : 0;
}
__device__ __forceinline__ double upw(double a, double b, double c, double d)
{
  return ((a - b) >= 0) ? c : d;
}

// Face fluxes with the operands and operation order of
// PL/nl_function_eval.c:696-854 (zero terrain slopes: x_dir_g = 0,
// x_dir_g_c = gravity, z_mult = FB = 1).  a is the lower-index cell.
__device__ __forceinline__ double flux_xy(double pa, double pb, double ka, double kb,
                                          double kra, double krb, double ff, double d,
                                          double gc, double mu)
{
  double diff = pa - pb;
  double updir = (diff / d) * gc;
  return ff * hmean(ka, kb) * (diff / d) * gc * upw(updir, 0.0, kra, krb) / mu;
}

__device__ __forceinline__ double flux_z(double pa, double pb, double ka, double kb,
                                         double kra, double krb, double ffz, double dz,
                                         double rho, double grav, double mu)
{
  double lc = pa / dz - 0.5 * rho * grav * 1.0;
  double uc = pb / dz + 0.5 * rho * grav * 1.0;
  double diff = lc - uc;
  return ffz * hmean_dz1(ka, kb) * diff * upw(lc, uc, kra, krb) / mu;
}

// F of one interior cell: accumulation + storage + sources (:394-540), the
// interior face fluxes (own faces out, incoming faces in, summed in the
// oracle's order) and the boundary-face terms (:975-2260).
// XY = false: the caller passes ul = 0 at i = 0 and ud = 0 at j = 0 (adding
// -dt * 0 leaves fv's bits unchanged), so the x and y tests are left out
template <bool XY = true>
__device__ __forceinline__ double residual_cell(const ResArgs &A, int i, int j, int k, int ip,
                                                double pc, double krc, double S,
                                                double kxc, double kyc, double kzc,
                                                double ur, double uf, double uu,
                                                double ul, double ud, double ub, double so,
                                                double po, double ssv, double pold)
{
  const Geom &g = A.g;
  const double rho = A.f.rho, grav = A.f.g, mu = A.f.mu;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const double vol = dx * dy * dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  double fv = (S * rho - so * rho) * po * vol;
  fv += ssv * vol * (pc * S * rho - pold * so * rho);
  if (A.src) fv -= vol * dt * (A.src[ip]);
  fv += dt * (ur + uf + uu);
  if (!XY || i > 0) fv += -dt * ul;
  if (!XY || j > 0) fv += -dt * ud;
  if (k > 0) fv += -dt * ub;
  return fv;
}

// DirichletBC "u_new" of face fc of a boundary cell (:1018-1232): the flux
// through the face from the Dirichlet value at half a cell's distance, with
// the ghost kr*rho taken at that value (inserted at :574-610).  F adds
// dt dir u_new (:1225); the face velocity is u_new / h (:1230).
__device__ __forceinline__ double dir_u_new(const ResArgs &A, int fc, int ip, double value,
                                            double pc, double krc, double kxc, double kyc,
                                            double kzc)
{
  const double rho = A.f.rho, grav = A.f.g, mu = A.f.mu;
  const double dx = A.dx, dy = A.dy, dz = A.dz;
  const double h = (fc < 2) ? dy * dz : (fc < 4) ? dx * dz : dx * dy;
  double kg = vg_kr(value, A.f.irg, A.f.alpha[ip], A.f.nvg[ip], A.f.tb) * rho;
  double u_new;
  if (fc == 0) {
    double diff = value - pc;
    u_new = upw(value, pc, kg, krc) * h * (kxc / mu) * 2.0 * (diff / dx);
  } else if (fc == 1) {
    double diff = pc - value;
    u_new = upw(pc, value, krc, kg) * h * (kxc / mu) * 2.0 * (diff / dx);
  } else if (fc == 2) {
    double diff = value - pc;
    u_new = upw(value, pc, kg, krc) * h * (kyc / mu) * 2.0 * (diff / dy);
  } else if (fc == 3) {
    double diff = pc - value;
    u_new = upw(pc, value, krc, kg) * h * (kyc / mu) * 2.0 * (diff / dy);
  } else if (fc == 4) {
    double sep = dz * 1.0;
    double lc = value / sep - 0.25 * rho * grav;
    double uc = pc / sep + 0.25 * rho * grav;
    double diff = lc - uc;
    u_new = upw(lc, uc, kg, krc) * h * (kzc / mu) * 2.0 * diff;
  } else {
    double sep = dz * 1.0;
    double lc = (pc / sep) - 0.25 * rho * grav * 1.0;
    double uc = (value / sep) + 0.25 * rho * grav * 1.0;
    double diff = lc - uc;
    u_new = upw(lc, uc, krc, kg) * h * (kzc / mu) * 2.0 * diff;
  }
  return u_new;
}

// The boundary-face terms of one cell (:975-2260), added to fv face by face
// in the order x-lower, x-upper, y-lower, y-upper, z-lower, z-upper, after
// all interior terms -- the order of the reference's patch loops.
__device__ __forceinline__ double residual_bc(const ResArgs &A, int i, int j, int k, int ip,
                                              double fv, double pc, double krc, double kxc,
                                              double kyc, double kzc)
{
  const Geom &g = A.g;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const double vol = dx * dy * dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  bool bnd = (i == 0) | (i == g.nx - 1) | (j == 0) | (j == g.ny - 1) | (k == 0) |
             (k == g.nz - 1);
  if (bnd) {
#pragma unroll 1
    for (int fc = 0; fc < 6; fc++) {
      bool on = (fc == 0 && i == 0) || (fc == 1 && i == g.nx - 1) ||
                (fc == 2 && j == 0) || (fc == 3 && j == g.ny - 1) ||
                (fc == 4 && k == 0) || (fc == 5 && k == g.nz - 1);
      if (!on) continue;
      int fi = (fc < 2) ? j + k * g.ny : (fc < 4) ? i + k * g.nx : i + j * g.nx;
      double value = A.bcv[fc][fi];
      int type = A.bct[fc];
      double dir = (fc & 1) ? 1.0 : -1.0;
      double h = (fc < 2) ? ffx : (fc < 4) ? ffy : ffz;
      if (type == PFHIP_BC_FLUX) {
        double u_new = h * value;
        fv += dt * dir * u_new;
      } else if (type == PFHIP_BC_DIR_EQUIL) {
        double u_new = dir_u_new(A, fc, ip, value, pc, krc, kxc, kyc, kzc);
        fv += dt * dir * u_new;
      } else if (fc == 5) {
        // OverlandKinematicBC, PL/nl_function_eval.c:1786-2010
        int io = (i + 1) + (j + 1) * g.sy;
        double ke = A.qx[io], kw = A.qx[io - 1];
        double kn = A.qy[io], ks = A.qy[io - g.sy];
        double q_overlnd = vol * (fmax(pc, 0.0) - fmax(A.p_old[ip], 0.0)) / dz +
                           dt * vol * ((ke - kw) / dx + (kn - ks) / dy) / dz;
        fv += q_overlnd;
        double u_new = h * value;
        fv += dt * dir * u_new;
      }
    }
  }
  return fv;
}

// Residual tiles are RTX columns wide (one wavefront per tile row).
#define RTX 64

// Face fluxes with reciprocal constants (1/d, 1/mu multiply instead of
// divide: within an ulp of the quotients) and the fast harmonic mean.  Same
// operand order as flux_xy / flux_z.
__device__ __forceinline__ double hmean_f(double a, double b)
{
  return (a + b) ? pf_div(2.0 * a * b, a + b) : 0;
}
__device__ __forceinline__ double flux_xy_i(double pa, double pb, double ka, double kb,
                                            double kra, double krb, double ff, double id,
                                            double gc, double imu)
{
  double dd = (pa - pb) * id;
  double updir = dd * gc;
  return ff * hmean_f(ka, kb) * dd * gc * upw(updir, 0.0, kra, krb) * imu;
}
__device__ __forceinline__ double flux_z_i(double pa, double pb, double ka, double kb,
                                           double kra, double krb, double ffz, double idz,
                                           double hrg, double imu)
{
  double lc = pa * idz - hrg;
  double uc = pb * idz + hrg;
  double diff = lc - uc;
  return ffz * hmean_f(ka, kb) * diff * upw(lc, uc, kra, krb) * imu;
}

// Software-pipelined fused residual.  A TX x TY block owns a tile of
// columns and marches up a z-chunk of ZC planes; each cell computes its
// closures one plane ahead and the fluxes through its own x+, y+, z+ faces
// once (the tile's west/south edge halo is recomputed); every global load of plane k+1
// (k+2 for the closure inputs) is issued at the top of plane k from a
// clamped in-bounds address, and the two per-plane barriers order LDS only,
// so the loads stay in flight while plane k is computed.  PCS: s_res/s_sat
// are per-cell fields.
// pr: the unshifted pressure (the line-search form's u, for KINStop's step)
template <bool PCS> struct VGIn { double p, pr, al, n, kz, sr, st; };
template <> struct VGIn<false> { double p, pr, al, n, kz; };
struct OthIn { double kx, kxe, ky, kyn, so, po, ss, pold, hk0, hk1, fs; };
struct HalIn { double p, al, n; };

// load through a 32-bit byte offset from the array base (every ghosted
// array is far below 4 GiB): the load then takes the base from scalar
// registers and one shared vector offset instead of a 64-bit address each
__device__ __forceinline__ double ldb(const double *__restrict__ base, int ip)
{
  return *(const double *)((const char *)base + ((unsigned)ip << 3));
}

// the pressure the residual is evaluated at: p, or p + psig pv
template <int MODE>
__device__ __forceinline__ double p_at(const ResArgs &A, int ip, double psig)
{
  if constexpr (MODE != 0) return ldb(A.p, ip) + psig * ldb(A.pv, ip);
  else return ldb(A.p, ip);
}

// the form's scalars: sigma and the output scale, from the device sums of
// the difference quotient when given (the expression of atimes in
// kinsol_host.c, KS/kinspgmr.c:500-515)
__device__ __forceinline__ void form_scalars(const ResArgs &A, double &psig, double &fsc)
{
  psig = A.psig;
  fsc = A.fsc;
  if (A.dq) {
    const double sutsv = A.dq[0], vtv = A.dq[1], sq1norm = A.dq[2];
    const double sign = (sutsv >= 0.0) ? 1.0 : -1.0;
    const double a = fabs(sutsv);
    psig = sign * A.sqrf * ((a > sq1norm) ? a : sq1norm) / vtv;
    fsc = 1.0 / psig;
  }
}

template <bool PCS, int MODE>
__device__ __forceinline__ void load_vg(VGIn<PCS> &v, const ResArgs &A, int ip, double psig)
{
  if constexpr (MODE == 2) {
    v.pr = ldb(A.p, ip);
    v.p = v.pr + psig * ldb(A.pv, ip);        // p_at's expression
  } else {
    v.p = p_at<MODE>(A, ip, psig);
  }
  v.al = ldb(A.f.alpha, ip); v.n = ldb(A.f.nvg, ip);
  v.kz = ldb(A.kz, ip);
  if constexpr (PCS) { v.sr = ldb(A.f.sres, ip); v.st = ldb(A.f.ssat, ip); }
}
template <bool PCS> __device__ __forceinline__ double vg_sr(const VGIn<PCS> &v, const Phys &f)
{
  if constexpr (PCS) return v.sr; else return f.sres_v;
}
template <bool PCS> __device__ __forceinline__ double vg_st(const VGIn<PCS> &v, const Phys &f)
{
  if constexpr (PCS) return v.st; else return f.ssat_v;
}
// ISO: Kx, Ky and Kz are one array (isotropic permeability); the cell's own
// value is then the Kz already carried for the z face
template <int MODE, bool ISO>
__device__ __forceinline__ void load_oth(OthIn &o, const ResArgs &A, int ip, int sy, int hip,
                                         const double *__restrict__ Kh, int hoff)
{
  if constexpr (!ISO) { o.kx = ldb(A.kx, ip); o.ky = ldb(A.ky, ip); }
  o.kxe = ldb(ISO ? A.kz : A.kx, ip + 1);
  o.kyn = ldb(ISO ? A.kz : A.ky, ip + sy);
  o.hk0 = ldb(Kh, hip); o.hk1 = ldb(Kh, hip + hoff);
  o.so = ldb(A.s_old, ip); o.po = ldb(A.poro, ip); o.ss = ldb(A.ss, ip);
  o.pold = ldb(A.p_old, ip);
  if constexpr (MODE == 1) o.fs = ldb(A.fsub, ip);
}
template <int MODE>
__device__ __forceinline__ void load_hal(HalIn &h, const ResArgs &A, int hip, double psig)
{
  h.p = p_at<MODE>(A, hip, psig); h.al = ldb(A.f.alpha, hip); h.n = ldb(A.f.nvg, hip);
}

template <int TX, int TY, int ZC, bool PCS, int MODE, bool ISO, bool TAB>
__global__ __launch_bounds__(TX * TY) void k_residual_pf(ResArgs A)
{
  constexpr int HALO = 2 * TX + 2 * TY;
  constexpr int W2 = TX + 2;
  constexpr int LDSN = (TX + 2) * (TY + 2);
  __shared__ double s_p[LDSN];
  __shared__ double s_kr[LDSN];
  __shared__ double s_ux[TY * (TX + 1)];   // x faces: west face of tile column tx at tx
  __shared__ double s_uy[(TY + 1) * TX];   // y faces: south face of tile row ty at ty
  const Geom g = A.g;
  const Phys &f = A.f;
  double psig = 0.0, fsc = 0.0;
  if constexpr (MODE != 0) form_scalars(A, psig, fsc);
  const double rho = f.rho, grav = f.g, irg = f.irg, imu = A.imu;
  const double hrg = 0.5 * rho * grav * 1.0;
  const double dx = A.dx, dy = A.dy, dz = A.dz;
  const double idx = A.idx, idy = A.idy, idz = A.idz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  int bx, by, bz;
  tile_of_block<MODE != 1>(bx, by, bz);
  const int i = bx * TX + tx, j = by * TY + ty;
  const bool act = (i < g.nx) && (j < g.ny);
  const int nz = g.nz, sy = g.sy, sz = g.sz;
  const int ls = (ty + 1) * W2 + (tx + 1);
  // loads use the clamped column; stores only happen for act cells
  const int col = (min(i, g.nx - 1) + 1) + (min(j, g.ny - 1) + 1) * sy;
  // edge-halo cell of this thread (corners are not needed by 7 points)
  int hx = 0, hy = 0;
  bool hact = false;
  if (tid < HALO) {
    if (tid < TX) { hx = tid + 1; hy = 0; }
    else if (tid < 2 * TX) { hx = tid - TX + 1; hy = TY + 1; }
    else if (tid < 2 * TX + TY) { hx = 0; hy = tid - 2 * TX + 1; }
    else { hx = TX + 1; hy = tid - 2 * TX - TY + 1; }
    int hi = bx * TX + hx - 1, hj = by * TY + hy - 1;
    hact = (hi >= 0 && hi < g.nx && hj >= 0 && hj < g.ny);
  }
  const int hls = hy * W2 + hx;
  // halo loads come from the halo cell when it exists, else the own column
  const int hcol = hact ? (bx * TX + hx) + (by * TY + hy) * sy : col;
  // west halo (hx == 0) supplies the x face into the tile, south halo
  // (hy == 0) the y face into the tile
  const bool hsouth = hact && hy == 0;
  // west / south halo threads, also where the halo cell is outside the domain
  const bool hw0 = tid < HALO && hx == 0, hs0 = tid < HALO && hy == 0;
  const double *__restrict__ Kh = ISO ? A.kz : hsouth ? A.ky : A.kx;
  const int hoff = hsouth ? sy : 1;
  const int k0 = bz * ZC;
  const int k1 = min(k0 + ZC, nz);
  auto plane = [&](int q) { return (min(q, nz - 1) + 1) * sz; };

  double p_m = 0, kr_m = 0, kz_m = 0;
  double p_c = 0, kr_c = 0, S_c = 0, kz_c = 0, pr_c = 0;
  double p_n = 0, kr_n = 0, S_n = 0, kz_n = 0, pr_n = 0;
  {
    VGIn<PCS> v0;
    load_vg<PCS, MODE>(v0, A, col + plane(k0), psig);
    if (k0 > 0) {
      VGIn<PCS> vm;
      load_vg<PCS, MODE>(vm, A, col + plane(k0 - 1), psig);
      p_m = vm.p;
      kr_m = vg_kr<TAB>(p_m, irg, vm.al, vm.n, f.tb) * rho;
      kz_m = vm.kz;
    }
    p_c = v0.p;
    if constexpr (MODE == 2) pr_c = v0.pr;
    vg_s_kr<TAB>(p_c, irg, v0.al, v0.n, vg_sr<PCS>(v0, f), vg_st<PCS>(v0, f), S_c, kr_c, f.tb);
    kr_c *= rho;
    kz_c = v0.kz;
  }
  double st_f = 0.0, st_s = 0.0;    // MODE 2 KINStop maxima of this thread
  // flux through the lower z face of the own cell at plane k0
  double ub = 0.0;
  if (k0 > 0) ub = flux_z_i(p_m, p_c, kz_m, kz_c, kr_m, kr_c, ffz, idz, hrg, imu);

  VGIn<PCS> vgN;
  HalIn hal;
  load_vg<PCS, MODE>(vgN, A, col + plane(k0 + 1), psig);
  load_hal<MODE>(hal, A, hcol + plane(k0), psig);

  for (int k = k0; k < k1; k++) {
    const int ip = col + (k + 1) * sz;
    // this plane's flux and storage inputs (used after the closures), then
    // the prefetch of the next plane's closure inputs
    OthIn o;
    load_oth<MODE, ISO>(o, A, col + plane(k), sy, hcol + plane(k), Kh, hoff);
    if constexpr (ISO) { o.kx = kz_c; o.ky = kz_c; }
    const VGIn<PCS> vgU = vgN;
    const HalIn h = hal;
    load_vg<PCS, MODE>(vgN, A, col + plane(k + 2), psig);
    load_hal<MODE>(hal, A, hcol + plane(k + 1), psig);
    // ---- A: closures of the own cell one plane ahead and of the halo cell
    if (k + 1 < nz) {
      p_n = vgU.p;
      if constexpr (MODE == 2) pr_n = vgU.pr;
      vg_s_kr<TAB>(p_n, irg, vgU.al, vgU.n, vg_sr<PCS>(vgU, f), vg_st<PCS>(vgU, f), S_n, kr_n, f.tb);
      kr_n *= rho;
      kz_n = vgU.kz;
    }
    if (hact) {
      s_p[hls] = h.p;
      s_kr[hls] = vg_kr<TAB>(h.p, irg, h.al, h.n, f.tb) * rho;
    }
    s_p[ls] = p_c;
    s_kr[ls] = kr_c;
    lds_barrier();
    // ---- B: own x+, y+, z+ faces; west / south halo faces into the tile
    double ur = 0.0, uf = 0.0, uu = 0.0;
    if (i + 1 < g.nx)
      ur = flux_xy_i(p_c, s_p[ls + 1], o.kx, o.kxe, kr_c, s_kr[ls + 1], ffx, idx, grav, imu);
    if (j + 1 < g.ny)
      uf = flux_xy_i(p_c, s_p[ls + W2], o.ky, o.kyn, kr_c, s_kr[ls + W2], ffy, idy, grav,
                     imu);
    if (k + 1 < nz) uu = flux_z_i(p_c, p_n, kz_c, kz_n, kr_c, kr_n, ffz, idz, hrg, imu);
    s_ux[ty * (TX + 1) + tx + 1] = ur;
    s_uy[(ty + 1) * TX + tx] = uf;
    // the west / south halo slots hold 0 where the tile is on the domain's
    // west / south side (no face there)
    if (hw0) {
      const double u = flux_xy_i(s_p[hls], s_p[hls + 1], o.hk0, o.hk1, s_kr[hls],
                                 s_kr[hls + 1], ffx, idx, grav, imu);
      s_ux[(hy - 1) * (TX + 1)] = hact ? u : 0.0;
    }
    if (hs0) {
      const double u = flux_xy_i(s_p[hls], s_p[hls + W2], o.hk0, o.hk1, s_kr[hls],
                                 s_kr[hls + W2], ffy, idy, grav, imu);
      s_uy[hx - 1] = hact ? u : 0.0;
    }
    lds_barrier();
    // ---- C
    if (act) {
      double ul = s_ux[ty * (TX + 1) + tx];
      double ud = s_uy[ty * TX + tx];
      double fv = residual_cell<false>(A, i, j, k, ip, p_c, kr_c, S_c, o.kx, o.ky, kz_c, ur, uf,
                                       uu, ul, ud, ub, o.so, o.po, o.ss, o.pold);
      if constexpr (MODE == 1) {
        // boundary cells get the transform after their face terms
        const bool bnd = (i == 0) | (i == g.nx - 1) | (j == 0) | (j == g.ny - 1) | (k == 0) |
                         (k == nz - 1);
        fv = bnd ? fv : fsc * fv - fsc * o.fs;
      }
      if constexpr (MODE == 2) {
        A.pout[ip] = p_c;
        // KINStop's terms (KS/kinsol.c:1507-1540) of the interior cells;
        // the boundary kernel takes the shell cells after their face terms
        if (A.stat) {
          const bool bnd = (i == 0) | (i == g.nx - 1) | (j == 0) | (j == g.ny - 1) | (k == 0) |
                           (k == nz - 1);
          if (!bnd) {
            st_f = fmax(st_f, fabs(fv));
            const double stp = 1.0 * p_c + (-1.0) * pr_c;
            st_s = fmax(st_s, fabs(stp / (1.0 + fabs(p_c))));
          }
        }
      }
      A.F[ip] = fv;
    }
    ub = uu;
    p_m = p_c; kr_m = kr_c; kz_m = kz_c;
    p_c = p_n; kr_c = kr_n; S_c = S_n; kz_c = kz_n; pr_c = pr_n;
  }
  if constexpr (MODE == 2) {
    if (A.stat) {
      block_max2<TX * TY>(st_f, st_s);
      if (tid == 0) {
        const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        A.stat[blk] = st_f;
        A.stat[A.stat_nb + blk] = st_s;
      }
    }
  }
}

// Boundary cells of the domain, each exactly once: both z planes first,
// then the y-lower/upper rows of the inner planes, then the x-lower/upper
// columns of the inner planes and rows.
__device__ __forceinline__ bool shell_cell(const Geom &g, long idx, int &i, int &j, int &k)
{
  const int nzb = (g.nz > 1) ? 2 : 1, nyb = (g.ny > 1) ? 2 : 1, nxb = (g.nx > 1) ? 2 : 1;
  const int nzi = max(g.nz - 2, 0), nyi = max(g.ny - 2, 0);
  const long nplane = (long)g.nx * g.ny;
  const long n0 = nplane * nzb, n1 = (long)g.nx * nyb * nzi, n2 = (long)nxb * nyi * nzi;
  if (idx < n0) {
    k = (idx >= nplane) ? g.nz - 1 : 0;
    long r = idx % nplane;
    i = (int)(r % g.nx); j = (int)(r / g.nx);
    return true;
  }
  idx -= n0;
  if (idx < n1) {
    i = (int)(idx % g.nx);
    long r = idx / g.nx;
    j = (r % nyb) ? g.ny - 1 : 0;
    k = 1 + (int)(r / nyb);
    return true;
  }
  idx -= n1;
  if (idx < n2) {
    i = (idx % nxb) ? g.nx - 1 : 0;
    long r = idx / nxb;
    j = 1 + (int)(r % nyi);
    k = 1 + (int)(r / nyi);
    return true;
  }
  return false;
}

static long shell_count(const Geom &g)
{
  const int nzb = (g.nz > 1) ? 2 : 1, nyb = (g.ny > 1) ? 2 : 1, nxb = (g.nx > 1) ? 2 : 1;
  const long nzi = max(g.nz - 2, 0), nyi = max(g.ny - 2, 0);
  return (long)g.nx * g.ny * nzb + (long)g.nx * nyb * nzi + (long)nxb * nyi * nzi;
}

// F += boundary-face terms (k_residual_pf leaves them out); in the J.v
// form the output transform follows them
template <int MODE>
__global__ __launch_bounds__(256) void k_residual_bc(ResArgs A, long nshell)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  double st_f = 0.0, st_s = 0.0;
  if (idx < nshell && shell_cell(A.g, idx, i, j, k)) {
    const int ip = (i + 1) + (j + 1) * A.g.sy + (k + 1) * A.g.sz;
    double psig = 0.0, fsc = 0.0;
    if constexpr (MODE != 0) form_scalars(A, psig, fsc);
    double pc = p_at<MODE>(A, ip, psig);
    double krc = vg_kr(pc, A.f.irg, A.f.alpha[ip], A.f.nvg[ip], A.f.tb) * A.f.rho;
    double fv = residual_bc(A, i, j, k, ip, A.F[ip], pc, krc, A.kx[ip], A.ky[ip], A.kz[ip]);
    if constexpr (MODE == 1) fv = fsc * fv - fsc * A.fsub[ip];
    A.F[ip] = fv;
    if constexpr (MODE == 2) {
      if (A.stat) {
        st_f = fabs(fv);
        const double stp = 1.0 * pc + (-1.0) * A.p[ip];
        st_s = fabs(stp / (1.0 + fabs(pc)));
      }
    }
  }
  if constexpr (MODE == 2) {
    if (A.stat) {
      block_max2<256>(st_f, st_s);
      if (threadIdx.x == 0) {
        A.stat[A.stat_main + blockIdx.x] = st_f;
        A.stat[A.stat_nb + A.stat_main + blockIdx.x] = st_s;
      }
    }
  }
}

// launch geometry of the tiled residual (launch_residual's 64 x 8 x 16
// tiles) and of the boundary-shell kernel
static long res_tiles(const Geom &g)
{
  return (long)((g.nx + 63) / 64) * ((g.ny + 7) / 8) * ((g.nz + 15) / 16);
}
static long res_bc_blocks(const Geom &g) { return (shell_count(g) + 255) / 256; }

template <int TX, int TY, int ZC, int MODE>
static void launch_pf(const ResArgs &A, hipStream_t st)
{
  const Geom &g = A.g;
  dim3 grid((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, (g.nz + ZC - 1) / ZC);
  dim3 block(TX, TY, 1);
  const int pid = (MODE == 1) ? PFHIP_PROF_RESIDUAL_JV
                  : (MODE == 2) ? PFHIP_PROF_RESIDUAL_STEP : PFHIP_PROF_RESIDUAL;
  PROF_BEGIN(pid, st);
  const bool iso = (A.kx == A.ky) && (A.kx == A.kz);
  if (A.f.tb.t) {
    // kr from a lookup table: the general (anisotropic) form
    if (A.f.sres || A.f.ssat)
      hipLaunchKernelGGL((k_residual_pf<TX, TY, ZC, true, MODE, false, true>), grid, block, 0, st, A);
    else
      hipLaunchKernelGGL((k_residual_pf<TX, TY, ZC, false, MODE, false, true>), grid, block, 0, st, A);
  } else if (A.f.sres || A.f.ssat) {
    if (iso) hipLaunchKernelGGL((k_residual_pf<TX, TY, ZC, true, MODE, true, false>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((k_residual_pf<TX, TY, ZC, true, MODE, false, false>), grid, block, 0, st, A);
  } else {
    if (iso) hipLaunchKernelGGL((k_residual_pf<TX, TY, ZC, false, MODE, true, false>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((k_residual_pf<TX, TY, ZC, false, MODE, false, false>), grid, block, 0, st, A);
  }
  PROF_END(pid, st);
  long ns = shell_count(g);
  hipLaunchKernelGGL(k_residual_bc<MODE>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                     st, A, ns);
}

static void launch_residual(const ResArgs &A, hipStream_t st)
{
  const int mode = (A.pv == nullptr) ? 0 : (A.fsub != nullptr) ? 1 : 2;
  if (mode == 1) launch_pf<64, 8, 16, 1>(A, st);
  else if (mode == 2) launch_pf<64, 8, 16, 2>(A, st);
  else launch_pf<64, 8, 16, 0>(A, st);
}

// the problem-fixed part of the residual arguments; the caller sets the
// pressure, output and form fields
static ResArgs res_args(const pfhip_problem *pb, const double *p_old, const double *s_old,
                        double dt)
{
  ResArgs A;
  A.g = geom_of(pb);
  A.f = phys_of(pb);
  A.dx = pb->d.dx; A.dy = pb->d.dy; A.dz = pb->d.dz; A.dt = dt;
  A.idx = 1.0 / A.dx; A.idy = 1.0 / A.dy; A.idz = 1.0 / A.dz; A.imu = 1.0 / A.f.mu;
  A.p_old = p_old; A.s_old = s_old;
  A.kx = pb->d.kx; A.ky = pb->d.ky; A.kz = pb->d.kz;
  A.poro = pb->d.poro; A.ss = pb->d.ss; A.src = pb->d.src;
  A.krho = pb->krho; A.qx = pb->q2[0]; A.qy = pb->q2[1];
  for (int f = 0; f < 6; f++) { A.bcv[f] = pb->bcv[f]; A.bct[f] = pb->d.bc[f].type; }
  A.p = nullptr; A.F = nullptr;
  A.pv = nullptr; A.fsub = nullptr; A.pout = nullptr; A.psig = 0.0; A.fsc = 0.0;
  A.dq = nullptr; A.sqrf = 0.0;
  A.stat = nullptr; A.stat_nb = 0; A.stat_main = 0;
  return A;
}

extern "C" int pfhip_nl_function_eval(pfhip_problem *pb, const double *p, double *F,
                                      const double *p_old, const double *s_old,
                                      double dt, void *stream)
{
  Geom g = geom_of(pb);
  if (pb->overland) {
    hipLaunchKernelGGL(k_overland_q, grid2(g), BLOCK3, 0, STREAM(stream), g,
                       surf_of(pb), p, pb->q2[0], pb->q2[1]);
    LAUNCHCHK();
  }
  ResArgs A = res_args(pb, p_old, s_old, dt);
  A.p = p;
  A.F = F;
  launch_residual(A, STREAM(stream));
  LAUNCHCHK();
  return PFHIP_OK;
}

// --------------------------------------------------------------------------
// face velocities -- the vx, vy, vz NlFunctionEval writes beside F
// (PL/nl_function_eval.c:846-848 interior faces; :1230 DirichletBC, :1432
// FluxBC, :2023 OverlandKinematicBC on domain faces), face-centred without
// ghosts, x fastest (the layout of the reference's velx/vely/velz output):
//   vx[i + (nx+1)(j + ny k)]  face i between cells i-1 and i, 0 <= i <= nx
//   vy[i + nx (j + (ny+1) k)] face j, 0 <= j <= ny
//   vz[i + nx (j + ny k)]     face k, 0 <= k <= nz
// Interior faces: the residual's face flux (same operands and order)
// divided by the face area ffx * z_mult * del_y_slope (= ffx here).
// One thread per point (i, j, k) of the (nx+1)(ny+1)(nz+1) lattice writes
// the x, y and z faces on the low side of cell (i, j, k); kr*rho comes from
// a closure pass into the problem's krho buffer.
// Algorithmic traffic per cell: p, kr*rho, Kx, Ky, Kz read (40 B) and three
// velocities written (24 B), plus the closure pass (p, alpha, n read, kr*rho
// written: 32 B).
// --------------------------------------------------------------------------
struct VelOut {
  double *__restrict__ vx;
  double *__restrict__ vy;
  double *__restrict__ vz;
};

// velocity the BC patch loop of face fc leaves at boundary cell (i, j, k)
__device__ __forceinline__ double bc_face_velocity(const ResArgs &A, int fc, int i, int j, int k)
{
  const Geom &g = A.g;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  const int fi = (fc < 2) ? j + k * g.ny : (fc < 4) ? i + k * g.nx : i + j * g.nx;
  const double value = A.bcv[fc][fi];
  const double h = (fc < 2) ? dy * dz : (fc < 4) ? dx * dz : dx * dy;
  const int type = A.bct[fc];
  if (type == PFHIP_BC_FLUX) return value;
  const double pc = A.p[ip];
  if (type == PFHIP_BC_DIR_EQUIL)
    return dir_u_new(A, fc, ip, value, pc, A.krho[ip], A.kx[ip], A.ky[ip], A.kz[ip]) / h;
  // OverlandKinematicBC (top face): u_new = h value + q_overlnd / dt
  const double vol = dx * dy * dz;
  const int io = (i + 1) + (j + 1) * g.sy;
  double q_overlnd = vol * (fmax(pc, 0.0) - fmax(A.p_old[ip], 0.0)) / dz +
                     dt * vol * ((A.qx[io] - A.qx[io - 1]) / dx +
                                 (A.qy[io] - A.qy[io - g.sy]) / dy) / dz;
  double u_new = h * value;
  u_new += q_overlnd / dt;
  return u_new / h;
}

__global__ __launch_bounds__(256) void k_face_velocity(ResArgs A, VelOut V)
{
  const Geom &g = A.g;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const double *__restrict__ p = A.p;
  const double *__restrict__ kr = A.krho;
  const double rho = A.f.rho, grav = A.f.g, mu = A.f.mu;
  const double dx = A.dx, dy = A.dy, dz = A.dz;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  // (i, j, k) may lie one past the last cell in any direction: ip is then
  // a ghost index, read only by faces that are interior
  const int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  if (j < g.ny && k < g.nz) {
    double v;
    if (i == 0) v = bc_face_velocity(A, 0, 0, j, k);
    else if (i == g.nx) v = bc_face_velocity(A, 1, g.nx - 1, j, k);
    else {
      const int a = ip - 1;
      v = flux_xy(p[a], p[ip], A.kx[a], A.kx[ip], kr[a], kr[ip], ffx, dx, grav, mu) /
          (ffx * 1.0 * 1.0);
    }
    V.vx[i + (long)(g.nx + 1) * (j + (long)g.ny * k)] = v;
  }
  if (i < g.nx && k < g.nz) {
    double v;
    if (j == 0) v = bc_face_velocity(A, 2, i, 0, k);
    else if (j == g.ny) v = bc_face_velocity(A, 3, i, g.ny - 1, k);
    else {
      const int a = ip - g.sy;
      v = flux_xy(p[a], p[ip], A.ky[a], A.ky[ip], kr[a], kr[ip], ffy, dy, grav, mu) /
          (ffy * 1.0 * 1.0);
    }
    V.vy[i + (long)g.nx * (j + (long)(g.ny + 1) * k)] = v;
  }
  if (i < g.nx && j < g.ny) {
    double v;
    if (k == 0) v = bc_face_velocity(A, 4, i, j, 0);
    else if (k == g.nz) v = bc_face_velocity(A, 5, i, j, g.nz - 1);
    else {
      const int a = ip - g.sz;
      v = flux_z(p[a], p[ip], A.kz[a], A.kz[ip], kr[a], kr[ip], ffz, dz, rho, grav, mu) /
          (ffz * 1.0 * 1.0);
    }
    V.vz[i + (long)g.nx * (j + (long)g.ny * k)] = v;
  }
}

extern "C" int pfhip_face_velocity(pfhip_problem *pb, const double *p, const double *p_old,
                                   double dt, double *vx, double *vy, double *vz, void *stream)
{
  if (!pb || !p || !vx || !vy || !vz) {
    pfhip_set_error("face_velocity: NULL argument");
    return PFHIP_ERR_ARG;
  }
  if (pb->overland && (!p_old || !(dt > 0.0))) {
    pfhip_set_error("face_velocity: an overland top needs p_old and dt > 0");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  hipStream_t st = STREAM(stream);
  if (pb->overland) {
    hipLaunchKernelGGL(k_overland_q, grid2(g), BLOCK3, 0, st, g, surf_of(pb), p, pb->q2[0],
                       pb->q2[1]);
    LAUNCHCHK();
  }
  int rc = launch_closure(pb, 4, p, pb->krho, nullptr, stream);
  if (rc) return rc;
  ResArgs A = res_args(pb, p_old, nullptr, dt);
  A.p = p;
  VelOut V;
  V.vx = vx; V.vy = vy; V.vz = vz;
  dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, g.nz + 1);
  hipLaunchKernelGGL(k_face_velocity, grid, dim3(64, 4, 1), 0, st, A, V);
  LAUNCHCHK();
  return PFHIP_OK;
}

// --------------------------------------------------------------------------
// water-balance terms of one step, volumes over the step (positive = water
// entering the domain, except the runoff, which leaves it):
//   [0] inflow through the domain faces other than an overland top:
//       -dt * sum dir h v over those faces, v the face velocity the BC
//       patch loops leave (PL/nl_function_eval.c:1230, :1432)
//   [1] inflow of an OverlandKinematic top's flux BC (rain): -dt h value
//   [2] surface runoff: dt * the kinematic fluxes leaving through the
//       domain's edges (the edge sum the q_overlnd divergence of
//       :1999-2004 telescopes to)
//   [3] sources: sum src vol dt (EvapTransSum's increment,
//       PL/evaptranssum.c:86-92)
// For a converged step, F = 0 summed over the cells gives
//   dM_sub + dM_surf = [0] + [1] - [2] + [3]
// with dM_sub the change of ParFlow's own stored mass (the residual's
// accumulation terms) and dM_surf that of the surface storage
// dx dy max(p_top, 0).  Row-owning blocks and a fixed-order final pass.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wb_terms(ResArgs A, int overland,
                                                  double *__restrict__ partial, int nb)
{
  const Geom &g = A.g;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const double vol_time = dx * dy * dz * dt;
  double bin = 0.0, rain = 0.0, run = 0.0, src = 0.0;
  const int rows = g.ny * g.nz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const int j = r % g.ny, k = r / g.ny;
    const int base = 1 + (j + 1) * g.sy + (k + 1) * g.sz;
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      const int ip = base + i;
      if (A.src) src += A.src[ip] * vol_time;
      const bool bnd = (i == 0) | (i == g.nx - 1) | (j == 0) | (j == g.ny - 1) | (k == 0) |
                       (k == g.nz - 1);
      if (!bnd) continue;
      for (int fc = 0; fc < 6; fc++) {
        const bool on = (fc == 0 && i == 0) || (fc == 1 && i == g.nx - 1) ||
                        (fc == 2 && j == 0) || (fc == 3 && j == g.ny - 1) ||
                        (fc == 4 && k == 0) || (fc == 5 && k == g.nz - 1);
        if (!on) continue;
        const double dir = (fc & 1) ? 1.0 : -1.0;
        const double h = (fc < 2) ? dy * dz : (fc < 4) ? dx * dz : dx * dy;
        if (fc == 5 && overland) {
          const int fi = i + j * g.nx;
          rain += -dt * (h * A.bcv[5][fi]);
          const int io = (i + 1) + (j + 1) * g.sy;
          if (i == g.nx - 1) run += dt * dy * A.qx[io];
          if (i == 0) run -= dt * dy * A.qx[io - 1];
          if (j == g.ny - 1) run += dt * dx * A.qy[io];
          if (j == 0) run -= dt * dx * A.qy[io - g.sy];
        } else {
          bin += -dt * dir * h * bc_face_velocity(A, fc, i, j, k);
        }
      }
    }
  }
  bin = block_reduce<false>(bin);
  rain = block_reduce<false>(rain);
  run = block_reduce<false>(run);
  src = block_reduce<false>(src);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = bin;
    partial[nb + blockIdx.x] = rain;
    partial[2 * nb + blockIdx.x] = run;
    partial[3 * nb + blockIdx.x] = src;
  }
}

extern "C" int pfhip_water_balance_fluxes(pfhip_problem *pb, const double *p,
                                          const double *p_old, double dt, double *out,
                                          void *stream)
{
  if (!pb || !p || !out || !(dt > 0.0)) {
    pfhip_set_error("water_balance_fluxes: NULL argument or dt <= 0");
    return PFHIP_ERR_ARG;
  }
  if (pb->overland && !p_old) {
    pfhip_set_error("water_balance_fluxes: an overland top needs p_old");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  hipStream_t st = STREAM(stream);
  if (pb->overland) {
    hipLaunchKernelGGL(k_overland_q, grid2(g), BLOCK3, 0, st, g, surf_of(pb), p, pb->q2[0],
                       pb->q2[1]);
    LAUNCHCHK();
  }
  int rc = launch_closure(pb, 4, p, pb->krho, nullptr, stream);
  if (rc) return rc;
  ResArgs A = res_args(pb, p_old, nullptr, dt);
  A.p = p;
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_wb_terms, dim3(nb), dim3(256), 0, st, A, pb->overland ? 1 : 0,
                     pb->red_partial, nb);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, st, pb->red_partial, nb,
                     PFHIP_WB_TERMS, pb->red_dev);
  LAUNCHCHK();
  rc = fetch_results(pb, PFHIP_WB_TERMS, stream);
  if (rc) return rc;
  for (int q = 0; q < PFHIP_WB_TERMS; q++) out[q] = pb->h_red[q];
  return PFHIP_OK;
}

// EvapTransSum (PL/evaptranssum.c:31-92): sum += src vol dt on every cell
__global__ __launch_bounds__(256) void k_evap_trans_sum(Geom g, double vol_time,
                                                        const double *__restrict__ src,
                                                        double *__restrict__ sum)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  sum[ip] += src[ip] * vol_time;
}

extern "C" int pfhip_evap_trans_sum(pfhip_problem *pb, double dt, double *sum, void *stream)
{
  if (!pb || !sum) {
    pfhip_set_error("evap_trans_sum: NULL argument");
    return PFHIP_ERR_ARG;
  }
  if (!pb->d.src) return PFHIP_OK;   // no source field: nothing to add
  Geom g = geom_of(pb);
  const double vol_time = pb->d.dx * pb->d.dy * pb->d.dz * dt;
  hipLaunchKernelGGL(k_evap_trans_sum, grid3(g), BLOCK3, 0, STREAM(stream), g, vol_time,
                     pb->d.src, sum);
  LAUNCHCHK();
  return PFHIP_OK;
}

// --------------------------------------------------------------------------
// surface predictor (PL/solver_richards.c:3236-3362): before a step's solve,
// each top cell compares the water it must take this step -- the source
// (evap_trans) volume, the Darcy inflow through its lower and lateral faces
// from face velocities, optionally the lateral surface term
// dt ((qx_i - qx_(i-1)) dy + (qy_j - qy_(j-1)) dx), added with the sign
// :3312-3321 give it -- with its
// free pore volume; if the inflow is larger and the cell is unsaturated
// (p < 0), its pressure becomes the ponding depth the excess implies (or
// the fixed PressureValue when that is > 0).  One thread per top cell.
// --------------------------------------------------------------------------
struct PredArgs {
  Geom g;
  double dx, dy, dz, dt, pv;
  int lateral;
  double *__restrict__ p;
  const double *__restrict__ sat;
  const double *__restrict__ poro;
  const double *__restrict__ src;
  const double *__restrict__ vx;
  const double *__restrict__ vy;
  const double *__restrict__ vz;
  const double *__restrict__ qx;
  const double *__restrict__ qy;
};

__global__ __launch_bounds__(256) void k_surface_predictor(PredArgs A)
{
  const Geom &g = A.g;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k = g.nz - 1;
  const int ip = (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const double po = A.poro[ip];
  const double vol = dx * dy * dz * 1.0 * po * A.sat[ip];
  const double flux_in = dx * dy * dz * 1.0 * (A.src ? A.src[ip] : 0.0) * dt;
  const double vol_max = dx * dy * dz * 1.0 * po;
  const long vxi = (i + 1) + (long)(g.nx + 1) * (j + (long)g.ny * k);
  const long vxi_im1 = vxi - 1;
  const long vyi = i + (long)g.nx * ((j + 1) + (long)(g.ny + 1) * k);
  const long vyi_jm1 = vyi - g.nx;
  const long vzi_km1 = i + (long)g.nx * (j + (long)g.ny * k);
  const double flux_darcy = A.vz[vzi_km1] * dx * dy * dt +
                            (-A.vx[vxi] + A.vx[vxi_im1]) * dy * dz * 1.0 * dt +
                            (-A.vy[vyi] + A.vy[vyi_jm1]) * dx * dz * 1.0 * dt;
  double flux_total;
  if (A.lateral) {
    double lat = 0.0;
    if (A.qx) {
      const int io = (i + 1) + (j + 1) * g.sy;
      const double qx_current = A.qx[io];
      const double qx_west = (i > 0) ? A.qx[io - 1] : 0.0;
      const double qy_current = A.qy[io];
      const double qy_south = (j > 0) ? A.qy[io - g.sy] : 0.0;
      lat = dt * ((qx_current - qx_west) * dy + (qy_current - qy_south) * dx);
    }
    flux_total = flux_in + flux_darcy + lat;
  } else {
    flux_total = flux_in + flux_darcy;
  }
  double press_pred = (flux_total - (vol_max - vol)) / (dx * dy * po);
  if (flux_total > (vol_max - vol)) {
    if (A.p[ip] < 0.0) {
      if (A.pv > 0.0) press_pred = A.pv;
      A.p[ip] = press_pred;
    }
  }
}

extern "C" int pfhip_surface_predictor(pfhip_problem *pb, double *p, const double *sat,
                                       const double *vx, const double *vy, const double *vz,
                                       const double *qx, const double *qy, double dt,
                                       double pressure_value, int lateral_flows,
                                       void *stream)
{
  if (!pb || !p || !sat || !vx || !vy || !vz || (!qx) != (!qy)) {
    pfhip_set_error("surface_predictor: NULL argument");
    return PFHIP_ERR_ARG;
  }
  PredArgs A;
  A.g = geom_of(pb);
  A.dx = pb->d.dx; A.dy = pb->d.dy; A.dz = pb->d.dz; A.dt = dt;
  A.pv = pressure_value; A.lateral = lateral_flows ? 1 : 0;
  A.p = p; A.sat = sat; A.poro = pb->d.poro; A.src = pb->d.src;
  A.vx = vx; A.vy = vy; A.vz = vz; A.qx = qx; A.qy = qy;
  hipLaunchKernelGGL(k_surface_predictor, grid2(A.g), BLOCK3, 0, STREAM(stream), A);
  LAUNCHCHK();
  return PFHIP_OK;
}

// Line-search trial point of KINLineSearch (KS/kinsol.c:1228-1405): unew =
// uu + rl pp formed inside the residual launch (written to unew) and F at it.
int pfhip_residual_step(pfhip_problem *pb, const double *uu, const double *pp, double rl,
                        double *unew, double *F, const double *p_old, const double *s_old,
                        double dt, void *stream)
{
  if (pb->overland) {
    pfhip_set_error("residual_step: overland problems take the two-pass form");
    return PFHIP_ERR_ARG;
  }
  ResArgs A = res_args(pb, p_old, s_old, dt);
  A.p = uu;
  A.F = F;
  A.pv = pp; A.pout = unew; A.psig = rl;
  launch_residual(A, STREAM(stream));
  LAUNCHCHK();
  return PFHIP_OK;
}

// The line-search point of pfhip_residual_step plus, for KINLineSearch and
// KINStop (KS/kinsol.c:1228-1405, 1507-1580), F.F (the pfhip_vec_dot pass,
// same partials and order) and the maxima max |F| and max |unew - u| /
// (1 + |unew|) that pfhip_kin_stop_norms computes, gathered by the residual
// launches themselves: one synchronisation instead of three.
int pfhip_residual_step_stats(pfhip_problem *pb, const double *uu, const double *pp, double rl,
                              double *unew, double *F, const double *p_old,
                              const double *s_old, double dt, double *ff, double *fmax,
                              double *steplength, void *stream)
{
  if (pb->overland) {
    pfhip_set_error("residual_step_stats: overland problems take the two-pass form");
    return PFHIP_ERR_ARG;
  }
  hipStream_t st = STREAM(stream);
  ResArgs A = res_args(pb, p_old, s_old, dt);
  A.p = uu;
  A.F = F;
  A.pv = pp; A.pout = unew; A.psig = rl;
  A.stat = pb->stat_partial;
  A.stat_main = res_tiles(A.g);
  A.stat_nb = pb->stat_nb;
  if (A.stat_main + res_bc_blocks(A.g) != A.stat_nb) {
    pfhip_set_error("residual_step_stats: partial table of %ld, launch needs %ld", A.stat_nb,
                    A.stat_main + res_bc_blocks(A.g));
    return PFHIP_ERR_ARG;
  }
  launch_residual(A, st);
  LAUNCHCHK();
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_reduce<RED_DOT>, dim3(nb), dim3(256), 0, st, g, F, F, pb->red_partial,
                     nb);
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, st, pb->red_partial, nb, 1,
                     pb->red_dev);
  hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(256), 0, st, pb->stat_partial,
                     (int)pb->stat_nb, 2, pb->red_dev + 1);
  LAUNCHCHK();
  int rc = fetch_results(pb, 3, stream);
  if (rc) return rc;
  *ff = pb->h_red[0];
  *fmax = pb->h_red[1];
  *steplength = pb->h_red[2];
  return PFHIP_OK;
}

// Difference-quotient J.v of KINSpgmrAtimesDQ (KS/kinspgmr.c:481-531) in one
// residual launch: z = sinv F(u + sigma v) - sinv fval, without the two
// separate vector passes.  Not for problems with overland flow (their
// surface fluxes are evaluated from the pressure beforehand).
int pfhip_residual_jv(pfhip_problem *pb, const double *u, const double *v, double sigma,
                      const double *fval, double sinv, double *z, const double *p_old,
                      const double *s_old, double dt, void *stream)
{
  if (pb->overland) {
    pfhip_set_error("residual_jv: overland problems take the two-pass form");
    return PFHIP_ERR_ARG;
  }
  ResArgs A = res_args(pb, p_old, s_old, dt);
  A.p = u;
  A.F = z;
  A.pv = v; A.fsub = fval; A.psig = sigma; A.fsc = sinv;
  launch_residual(A, STREAM(stream));
  LAUNCHCHK();
  return PFHIP_OK;
}

// The same J.v with sigma taken on the device: the DQ sums (u.v, v.v,
// |v|_1) stay in the problem's result slots and the residual kernel forms
// sigma and 1/sigma from them, so no host synchronisation separates the
// reduction from the residual.
// the difference-quotient sums (u.v, v.v, |v|_1) the problem's DQ slots
// hold, to the host (after the work queued on the stream)
int pfhip_dq_slots(pfhip_problem *pb, double *out, void *stream)
{
  HIPCHK(hipMemcpyAsync(out, pb->red_dev + PFHIP_SLOT_DQ, 3 * sizeof(double),
                        hipMemcpyDeviceToHost, STREAM(stream)));
  HIPCHK(hipStreamSynchronize(STREAM(stream)));
  return PFHIP_OK;
}

// the difference-quotient J.v with its sums u.v, v.v and |v|_1 already in
// the problem's DQ slots (gathered by the V-cycle that produced v,
// pfhip_mg_apply_dq): the residual launch alone
int pfhip_residual_jv_ready(pfhip_problem *pb, const double *u, const double *v, double sqrf,
                            const double *fval, double *z, const double *p_old,
                            const double *s_old, double dt, void *stream)
{
  if (pb->overland) {
    pfhip_set_error("residual_jv: overland problems take the two-pass form");
    return PFHIP_ERR_ARG;
  }
  ResArgs A = res_args(pb, p_old, s_old, dt);
  A.p = u;
  A.F = z;
  A.pv = v; A.fsub = fval;
  A.dq = pb->red_dev + PFHIP_SLOT_DQ; A.sqrf = sqrf;
  launch_residual(A, STREAM(stream));
  LAUNCHCHK();
  return PFHIP_OK;
}

int pfhip_residual_jv_dq(pfhip_problem *pb, const double *u, const double *v, double sqrf,
                         const double *fval, double *z, const double *p_old,
                         const double *s_old, double dt, void *stream)
{
  if (pb->overland) {
    pfhip_set_error("residual_jv: overland problems take the two-pass form");
    return PFHIP_ERR_ARG;
  }
  Geom g = geom_of(pb);
  int nb = red_blocks(g);
  hipLaunchKernelGGL(k_reduce<RED_DQ>, dim3(nb), dim3(256), 0, STREAM(stream), g, u, v,
                     pb->red_partial, nb);
  hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(256), 0, STREAM(stream),
                     pb->red_partial, nb, 3, pb->red_dev + PFHIP_SLOT_DQ);
  LAUNCHCHK();
  ResArgs A = res_args(pb, p_old, s_old, dt);
  A.p = u;
  A.F = z;
  A.pv = v; A.fsub = fval;
  A.dq = pb->red_dev + PFHIP_SLOT_DQ; A.sqrf = sqrf;
  launch_residual(A, STREAM(stream));
  LAUNCHCHK();
  return PFHIP_OK;
}

// --------------------------------------------------------------------------
// Jacobian -- RichardsJacobianEval, PL/richards_jacobian_eval.c:217-2200,
// gather form: each cell writes its own row.  PFSymmetric (symm_part)
// keeps the full derivative on the diagonal and the symmetric part off the
// diagonal (:594-776); Dirichlet faces :1049-1226; kinematic overland top
// rows :1547-1644/:1938-1987 with the PFSymmetric top-row assembly of
// PL/pf_hypre.c:360-420.
// --------------------------------------------------------------------------
struct JacArgs {
  Geom g;
  Phys f;
  double dx, dy, dz, dt;
  const double *__restrict__ p;
  const double *__restrict__ kx;
  const double *__restrict__ ky;
  const double *__restrict__ kz;
  const double *__restrict__ poro;
  const double *__restrict__ ss;
  const double *__restrict__ krho;
  const double *__restrict__ dkrho;
  const double *__restrict__ ke;
  const double *__restrict__ kw;
  const double *__restrict__ kn;
  const double *__restrict__ ks;
  const double *bcv[6];
  int bct[6];
  int symm;
  int overland;
  double *J[7];
};

// ---- fused, tiled Jacobian -------------------------------------------------
// One launch per assembly: a TX x TY tile of columns marches a z-chunk like
// k_residual_pf; every cell evaluates ONE van Genuchten chain (S, dS/dp, kr,
// dkr/dp), every face is evaluated once by its lower cell, which keeps its
// own terms and passes the upper cell's through LDS (x, y faces) or a
// register (z face).  Per face (a = lower cell, b = upper) the reference's
// two branches (PL/richards_jacobian_eval.c:594-776) reduce to
//   T1 = -(a's derivative term) + sym   ->  c_a -= T1, full-J W/S/L of b
//   T2 =  (b's derivative term) + sym   ->  c_b -= T2, full-J E/N/U of a
// with the symmetric part sym stored off the diagonal by PFSymmetric.  The
// diagonal subtracts its faces in the reference's order E, W, N, S, U, L.
// Dirichlet faces and the overland top rows are applied afterwards by
// k_jacobian_bc.  OT is the coefficient type (float: the FP32 multigrid's
// finest-level matrix is written directly).
template <typename OT> struct JOut { OT *c[7]; };

struct JacPFArgs {
  JacArgs a;
  double idx, idy, idz, imu;
};

template <bool TAB = true>
__device__ __forceinline__ void vg_all(double p, double irg, double alpha, double n,
                                       double sres, double ssat, double &S, double &dS,
                                       double &kr, double &dkr, const KrTab &tb)
{
  double s_dif = ssat - sres;
  if (p >= 0.0) { S = s_dif + sres; dS = 0.0; kr = 1.0; dkr = 0.0; return; }
  VGTerms v = vg_terms(p, irg, alpha, n);
  S = s_dif * v.a + sres;
  dS = (v.m * n * alpha * v.ahnm1) * s_dif * pf_div(v.a, 1.0 + v.t);
  if (TAB && tb.t) {
    const double head = fabs(p) * irg;
    kr = kr_table(tb, head, false);
    dkr = kr_table(tb, head, true);
  } else {
    kr = vg_kr_t(v);
    dkr = vg_dkr_t(v, alpha, n);
  }
}

template <bool PCS> struct JVG { double p, al, n, kz, sr, st; };
template <> struct JVG<false> { double p, al, n, kz; };
struct JOth { double kx, kxe, ky, kyn, po, ss, hk0, hk1; };
struct JHal { double p, al, n; };

template <bool PCS>
__device__ __forceinline__ void jload_vg(JVG<PCS> &v, const JacArgs &A, int ip)
{
  v.p = A.p[ip]; v.al = A.f.alpha[ip]; v.n = A.f.nvg[ip]; v.kz = A.kz[ip];
  if constexpr (PCS) { v.sr = A.f.sres[ip]; v.st = A.f.ssat[ip]; }
}
template <bool PCS> __device__ __forceinline__ double jsr(const JVG<PCS> &v, const Phys &f)
{
  if constexpr (PCS) return v.sr; else return f.sres_v;
}
template <bool PCS> __device__ __forceinline__ double jst(const JVG<PCS> &v, const Phys &f)
{
  if constexpr (PCS) return v.st; else return f.ssat_v;
}

// x or y face: T1 (for c_a, full-J W/S of b), T2 (for c_b, full-J E/N of a)
// and the symmetric couplings; ysign selects the y-face spelling whose
// symmetric part is written yc * -upw (same value)
__device__ __forceinline__ void jface_xy(double pa, double pb, double ka, double kb,
                                         double kra, double krb, double dka, double dkb,
                                         double dt, double ff, double id, double gc,
                                         double imu, double &T1, double &T2, double &sym)
{
  double diff = pa - pb;
  double updir = (diff * id) * gc;
  double xc = 1.0 * dt * ff * id * 1.0 * hmean_f(ka, kb) * imu;
  sym = (-xc * upw(updir, 0.0, kra, krb)) * gc;
  T1 = (-xc * diff * upw(updir, 0.0, dka, 0.0)) * gc + sym;
  T2 = (xc * diff * upw(updir, 0.0, 0.0, dkb)) * gc + sym;
}

// z face (constant density: the density-derivative terms vanish)
__device__ __forceinline__ void jface_z(double pa, double pb, double ka, double kb,
                                        double kra, double krb, double dka, double dkb,
                                        double dt, double ffz, double isep, double hrg,
                                        double imu, double &T1, double &T2, double &sym_l,
                                        double &sym_u)
{
  double lc = pa * isep - hrg;
  double uc = pb * isep + hrg;
  double diff = lc - uc;
  double zc = 1.0 * dt * ffz * hmean_f(ka, kb) * imu;
  double up = upw(lc, uc, kra, krb);
  sym_l = -zc * isep * up;
  sym_u = zc * isep * -up;
  T1 = -zc * (diff * upw(lc, uc, dka, 0.0)) + sym_l;
  T2 = zc * (diff * upw(lc, uc, 0.0, dkb)) + sym_u;
}

template <int TX, int TY, int ZC, bool PCS, bool SYMM, typename OT, bool TAB>
// four waves per SIMD: the register budget of 128 VGPRs (without the hint
// the compiler took 131 and three waves; 401 -> 328 us on 16.8M cells)
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(4)))
void k_jacobian_pf(JacPFArgs P, JOut<OT> J)
{
  constexpr int HALO = 2 * TX + 2 * TY;
  constexpr int W2 = TX + 2;
  constexpr int LDSN = (TX + 2) * (TY + 2);
  __shared__ double s_p[LDSN];
  __shared__ double s_kr[LDSN];
  __shared__ double s_dk[LDSN];
  __shared__ double s_xc[TY * (TX + 1)];   // x faces: T2 for the east cell
  __shared__ double s_yc[(TY + 1) * TX];   // y faces: T2 for the north cell
  __shared__ double s_xw[SYMM ? 1 : TY * (TX + 1)];   // full J: W of the east cell
  __shared__ double s_ys[SYMM ? 1 : (TY + 1) * TX];   // full J: S of the north cell
  const JacArgs &A = P.a;
  const Geom g = A.g;
  const Phys &f = A.f;
  const double rho = f.rho, gc = f.g, irg = f.irg, imu = P.imu;
  const double dt = A.dt, dx = A.dx, dy = A.dy, dz = A.dz;
  const double idx = P.idx, idy = P.idy;
  const double isep = 1.0 / (dz * 0.5 * (1.0 + 1.0));
  const double hrg = (1.0 / (1.0 + 1.0)) * rho * gc;
  const double vol2 = dx * dy * dz * 1.0;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double ddp = 0.0;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  int bx, by, bz;
  tile_of_block(bx, by, bz);
  const int i = bx * TX + tx, j = by * TY + ty;
  const bool act = (i < g.nx) && (j < g.ny);
  const int nz = g.nz, sy = g.sy, sz = g.sz;
  const int ls = (ty + 1) * W2 + (tx + 1);
  const int col = (min(i, g.nx - 1) + 1) + (min(j, g.ny - 1) + 1) * sy;
  int hx = 0, hy = 0;
  bool hact = false;
  if (tid < HALO) {
    if (tid < TX) { hx = tid + 1; hy = 0; }
    else if (tid < 2 * TX) { hx = tid - TX + 1; hy = TY + 1; }
    else if (tid < 2 * TX + TY) { hx = 0; hy = tid - 2 * TX + 1; }
    else { hx = TX + 1; hy = tid - 2 * TX - TY + 1; }
    int hi = bx * TX + hx - 1, hj = by * TY + hy - 1;
    hact = (hi >= 0 && hi < g.nx && hj >= 0 && hj < g.ny);
  }
  const int hls = hy * W2 + hx;
  const int hcol = hact ? (bx * TX + hx) + (by * TY + hy) * sy : col;
  const bool hwest = hact && hx == 0;
  const bool hsouth = hact && hy == 0;
  const double *__restrict__ Kh = hsouth ? A.ky : A.kx;
  const int hoff = hsouth ? sy : 1;
  const int k0 = bz * ZC;
  const int k1 = min(k0 + ZC, nz);
  auto plane = [&](int q) { return (min(q, nz - 1) + 1) * sz; };

  double p_c = 0, S_c = 0, dS_c = 0, kr_c = 0, dk_c = 0, kz_c = 0;
  double p_n = 0, S_n = 0, dS_n = 0, kr_n = 0, dk_n = 0, kz_n = 0;
  double T2b = 0.0, T1b = 0.0, syml_b = 0.0;   // z face below the own cell
  {
    JVG<PCS> v0;
    jload_vg<PCS>(v0, A, col + plane(k0));
    p_c = v0.p;
    vg_all<TAB>(p_c, irg, v0.al, v0.n, jsr<PCS>(v0, f), jst<PCS>(v0, f), S_c, dS_c, kr_c, dk_c, f.tb);
    kr_c *= rho;
    dk_c = dk_c * rho + kr_c * 0.0;
    kz_c = v0.kz;
    if (k0 > 0) {
      JVG<PCS> vm;
      jload_vg<PCS>(vm, A, col + plane(k0 - 1));
      double Sm, dSm, krm, dkm;
      vg_all<TAB>(vm.p, irg, vm.al, vm.n, jsr<PCS>(vm, f), jst<PCS>(vm, f), Sm, dSm, krm, dkm, f.tb);
      krm *= rho;
      dkm = dkm * rho + krm * 0.0;
      double symu;
      jface_z(vm.p, p_c, vm.kz, kz_c, krm, kr_c, dkm, dk_c, dt, ffz, isep, hrg, imu, T1b,
              T2b, syml_b, symu);
    }
  }
  JVG<PCS> vgN;
  JHal hal;
  jload_vg<PCS>(vgN, A, col + plane(k0 + 1));
  hal.p = A.p[hcol + plane(k0)]; hal.al = f.alpha[hcol + plane(k0)];
  hal.n = f.nvg[hcol + plane(k0)];

  for (int k = k0; k < k1; k++) {
    const int ip = col + (k + 1) * sz;
    JOth o;
    {
      const int q = col + plane(k), hq = hcol + plane(k);
      o.kx = A.kx[q]; o.kxe = A.kx[q + 1]; o.ky = A.ky[q]; o.kyn = A.ky[q + sy];
      o.po = A.poro[q]; o.ss = A.ss[q];
      o.hk0 = Kh[hq]; o.hk1 = Kh[hq + hoff];
    }
    const JVG<PCS> vgU = vgN;
    const JHal h = hal;
    jload_vg<PCS>(vgN, A, col + plane(k + 2));
    {
      const int hq = hcol + plane(k + 1);
      hal.p = A.p[hq]; hal.al = f.alpha[hq]; hal.n = f.nvg[hq];
    }
    // ---- A: closures one plane ahead (own) and of the halo cell
    if (k + 1 < nz) {
      p_n = vgU.p;
      vg_all<TAB>(p_n, irg, vgU.al, vgU.n, jsr<PCS>(vgU, f), jst<PCS>(vgU, f), S_n, dS_n, kr_n,
             dk_n, f.tb);
      kr_n *= rho;
      dk_n = dk_n * rho + kr_n * 0.0;
      kz_n = vgU.kz;
    }
    if (hact) {
      double Sh, dSh, krh, dkh;
      vg_all<TAB>(h.p, irg, h.al, h.n, f.sres_v, f.ssat_v, Sh, dSh, krh, dkh, f.tb);
      krh *= rho;
      s_p[hls] = h.p;
      s_kr[hls] = krh;
      s_dk[hls] = dkh * rho + krh * 0.0;
    }
    s_p[ls] = p_c;
    s_kr[ls] = kr_c;
    s_dk[ls] = dk_c;
    lds_barrier();
    // ---- B: own E, N, U faces; west / south halo faces into the tile
    double T1e = 0, T1n = 0, T1u = 0, T2u = 0, offE = 0, offN = 0, offU = 0, symlu = 0;
    if (i + 1 < g.nx) {
      double T2, sym;
      jface_xy(p_c, s_p[ls + 1], o.kx, o.kxe, kr_c, s_kr[ls + 1], dk_c, s_dk[ls + 1], dt, ffx,
               idx, gc, imu, T1e, T2, sym);
      offE = SYMM ? sym : T2;
      s_xc[ty * (TX + 1) + tx + 1] = T2;
      if constexpr (!SYMM) s_xw[ty * (TX + 1) + tx + 1] = T1e;
    }
    if (j + 1 < g.ny) {
      double T2, sym;
      jface_xy(p_c, s_p[ls + W2], o.ky, o.kyn, kr_c, s_kr[ls + W2], dk_c, s_dk[ls + W2], dt,
               ffy, idy, gc, imu, T1n, T2, sym);
      offN = SYMM ? sym : T2;
      s_yc[(ty + 1) * TX + tx] = T2;
      if constexpr (!SYMM) s_ys[(ty + 1) * TX + tx] = T1n;
    }
    if (k + 1 < nz) {
      double syml, symu;
      jface_z(p_c, p_n, kz_c, kz_n, kr_c, kr_n, dk_c, dk_n, dt, ffz, isep, hrg, imu, T1u, T2u,
              syml, symu);
      offU = SYMM ? symu : T2u;
      symlu = syml;
    }
    if (hwest) {
      double T1, T2, sym;
      jface_xy(s_p[hls], s_p[hls + 1], o.hk0, o.hk1, s_kr[hls], s_kr[hls + 1], s_dk[hls],
               s_dk[hls + 1], dt, ffx, idx, gc, imu, T1, T2, sym);
      s_xc[(hy - 1) * (TX + 1)] = T2;
      if constexpr (!SYMM) s_xw[(hy - 1) * (TX + 1)] = T1;
    }
    if (hsouth) {
      double T1, T2, sym;
      jface_xy(s_p[hls], s_p[hls + W2], o.hk0, o.hk1, s_kr[hls], s_kr[hls + W2], s_dk[hls],
               s_dk[hls + W2], dt, ffy, idy, gc, imu, T1, T2, sym);
      s_yc[hx - 1] = T2;
      if constexpr (!SYMM) s_ys[hx - 1] = T1;
    }
    lds_barrier();
    // ---- C: the row of the own cell
    if (act) {
      double c = (dS_c * rho + S_c * ddp) * o.po * vol2 +
                 o.ss * vol2 * (dS_c * rho * p_c + S_c * ddp * p_c + S_c * rho);
      if (i + 1 < g.nx) c -= T1e;
      if (i > 0) c -= s_xc[ty * (TX + 1) + tx];
      if (j + 1 < g.ny) c -= T1n;
      if (j > 0) c -= s_yc[ty * TX + tx];
      if (k + 1 < nz) c -= T1u;
      if (k > 0) c -= T2b;
      J.c[0][ip] = (OT)c;
      J.c[2][ip] = (OT)offE;
      J.c[4][ip] = (OT)offN;
      J.c[6][ip] = (OT)offU;
      if constexpr (!SYMM) {
        J.c[1][ip] = (OT)((i > 0) ? s_xw[ty * (TX + 1) + tx] : 0.0);
        J.c[3][ip] = (OT)((j > 0) ? s_ys[ty * TX + tx] : 0.0);
        J.c[5][ip] = (OT)((k > 0) ? T1b : 0.0);
      }
    }
    // the z face above becomes the face below
    T2b = T2u; T1b = T1u; syml_b = symlu;
    p_c = p_n; S_c = S_n; dS_c = dS_n; kr_c = kr_n; dk_c = dk_n; kz_c = kz_n;
  }
}

// Dirichlet faces of one shell cell's row (PL/richards_jacobian_eval.c:
// 1049-1226), subtracted from the diagonal c face by face in face order
__device__ __forceinline__ double jac_dir_faces(const JacArgs &A, int i, int j, int k, int ip,
                                                double c)
{
  const Geom &g = A.g;
  const double rho = A.f.rho, grav = A.f.g, mu = A.f.mu;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double ddp = 0.0;
  const double pc = A.p[ip];
  const double a = A.f.alpha[ip], nn = A.f.nvg[ip];
  double S, dS, kr, dkr;
  vg_all(pc, A.f.irg, a, nn, A.f.Sres(ip), A.f.Ssat(ip), S, dS, kr, dkr, A.f.tb);
  const double krc = kr * rho, dkc = dkr * rho + krc * 0.0;
#pragma unroll 1
  for (int fc = 0; fc < 6; fc++) {
    bool on = (fc == 0 && i == 0) || (fc == 1 && i == g.nx - 1) ||
              (fc == 2 && j == 0) || (fc == 3 && j == g.ny - 1) ||
              (fc == 4 && k == 0) || (fc == 5 && k == g.nz - 1);
    if (!on || A.bct[fc] != PFHIP_BC_DIR_EQUIL) continue;
    int fi = (fc < 2) ? j + k * g.ny : (fc < 4) ? i + k * g.nx : i + j * g.nx;
    double value = A.bcv[fc][fi];
    double den_d = rho;
    double prod = krc, prod_der = dkc;
    double prod_val = vg_kr(value, A.f.irg, a, nn, A.f.tb) * den_d;
    double coeff, diff, o_temp;
    if (fc == 0 || fc == 2) {
      double ff = (fc == 0) ? ffx : ffy, dd = (fc == 0) ? dx : dy;
      double kk = (fc == 0) ? A.kx[ip] : A.ky[ip];
      coeff = dt * ff * 1.0 * (2.0 / dd) * kk / mu;
      diff = value - pc;
      o_temp = coeff * (diff * upw(value, pc, 0.0, prod_der) - upw(value, pc, prod_val, prod));
    } else if (fc == 1 || fc == 3) {
      double ff = (fc == 1) ? ffx : ffy, dd = (fc == 1) ? dx : dy;
      double kk = (fc == 1) ? A.kx[ip] : A.ky[ip];
      coeff = dt * ff * 1.0 * (2.0 / dd) * kk / mu;
      diff = pc - value;
      o_temp = -coeff * (diff * upw(pc, value, prod_der, 0.0) + upw(pc, value, prod, prod_val));
    } else if (fc == 4) {
      coeff = dt * ffz * (2.0 / (dz * 0.5 * (1.0 + 1.0))) * A.kz[ip] / mu;
      double lc = (value) - 0.5 * dz * 1.0 * den_d * grav;
      double uc = (pc) + 0.5 * dz * 1.0 * rho * grav;
      diff = lc - uc;
      o_temp = coeff * (diff * upw(lc, uc, 0.0, prod_der) +
                        ((-1.0 - grav * 0.5 * dz * 1.0 * ddp) * upw(lc, uc, prod_val, prod)));
    } else {
      coeff = dt * ffz * (2.0 / (dz * 0.5 * (1.0 + 1.0))) * A.kz[ip] / mu;
      double lc = (pc) - 0.5 * dz * 1.0 * rho * grav;
      double uc = (value) + 0.5 * dz * 1.0 * den_d * grav;
      diff = lc - uc;
      o_temp = -coeff * (diff * upw(lc, uc, prod_der, 0.0) +
                         ((1.0 - grav * 0.5 * dz * 1.0 * ddp) * upw(lc, uc, prod, prod_val)));
    }
    c -= o_temp;
  }
  return c;
}

// Dirichlet faces (:1049-1226) and kinematic overland top rows
// (:1547-1644/:1938-1987, PFSymmetric top rows of PL/pf_hypre.c:360-420) of
// the shell cells, applied to the assembled rows in the reference's order.
template <bool SYMM, typename OT>
__global__ __launch_bounds__(256) void k_jacobian_bc(JacArgs A, JOut<OT> J, long nshell)
{
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nshell) return;
  int i, j, k;
  const Geom &g = A.g;
  if (!shell_cell(g, idx, i, j, k)) return;
  const bool dir = (i == 0 && A.bct[0] == PFHIP_BC_DIR_EQUIL) ||
                   (i == g.nx - 1 && A.bct[1] == PFHIP_BC_DIR_EQUIL) ||
                   (j == 0 && A.bct[2] == PFHIP_BC_DIR_EQUIL) ||
                   (j == g.ny - 1 && A.bct[3] == PFHIP_BC_DIR_EQUIL) ||
                   (k == 0 && A.bct[4] == PFHIP_BC_DIR_EQUIL) ||
                   (k == g.nz - 1 && A.bct[5] == PFHIP_BC_DIR_EQUIL);
  const bool ovl = A.overland && k == g.nz - 1;
  if (!dir && !ovl) return;
  const int sy = g.sy;
  const int ip = (i + 1) + (j + 1) * sy + (k + 1) * g.sz;
  const double dx = A.dx, dy = A.dy, dz = A.dz, dt = A.dt;
  const double vol = dx * dy * dz;
  const double ffx = dy * dz, ffy = dx * dz;
  double c = (double)J.c[0][ip];
  const double pc = A.p[ip];
  if (dir) c = jac_dir_faces(A, i, j, k, ip, c);
  if (ovl) {
    int io = (i + 1) + (j + 1) * sy;
    double dvol = 0.0;
    if (pc > 0.0)
      dvol = (vol / dz) + (vol / ffy) * dt * (A.ke[io] - A.kw[io]) +
             (vol / ffx) * dt * (A.kn[io] - A.ks[io]);
    c += dvol;
    if (SYMM) {
      if (i + 1 < g.nx) J.c[2][ip] = (OT)0.0;
      if (j + 1 < g.ny) J.c[4][ip] = (OT)0.0;
    } else {
      double offW = (double)J.c[1][ip], offE = (double)J.c[2][ip];
      double offS = (double)J.c[3][ip], offN = (double)J.c[4][ip];
      offW -= (vol / ffy) * dt * (A.ke[io - 1]);
      offE += (vol / ffy) * dt * (A.kw[io + 1]);
      offS -= (vol / ffx) * dt * (A.kn[io - sy]);
      offN += (vol / ffx) * dt * (A.ks[io + sy]);
      if (i == 0) offW = 0.0;
      if (i == g.nx - 1) offE = 0.0;
      if (j == 0) offS = 0.0;
      if (j == g.ny - 1) offN = 0.0;
      J.c[1][ip] = (OT)offW; J.c[2][ip] = (OT)offE;
      J.c[3][ip] = (OT)offS; J.c[4][ip] = (OT)offN;
    }
  }
  J.c[0][ip] = (OT)c;
}

template <bool SYMM, typename OT>
static void launch_jacobian_pf(const JacPFArgs &P, const JOut<OT> &J, hipStream_t st)
{
  const Geom &g = P.a.g;
  constexpr int TX = 64, TY = 8, ZC = 16;
  dim3 grid((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, (g.nz + ZC - 1) / ZC);
  dim3 block(TX, TY, 1);
  PROF_BEGIN(PFHIP_PROF_JACOBIAN, st);
  const bool tab = P.a.f.tb.t != nullptr;
  if (P.a.f.sres || P.a.f.ssat) {
    if (tab) hipLaunchKernelGGL((k_jacobian_pf<TX, TY, ZC, true, SYMM, OT, true>), grid, block, 0, st, P, J);
    else hipLaunchKernelGGL((k_jacobian_pf<TX, TY, ZC, true, SYMM, OT, false>), grid, block, 0, st, P, J);
  } else {
    if (tab) hipLaunchKernelGGL((k_jacobian_pf<TX, TY, ZC, false, SYMM, OT, true>), grid, block, 0, st, P, J);
    else hipLaunchKernelGGL((k_jacobian_pf<TX, TY, ZC, false, SYMM, OT, false>), grid, block, 0, st, P, J);
  }
  PROF_END(PFHIP_PROF_JACOBIAN, st);
  long ns = shell_count(g);
  hipLaunchKernelGGL((k_jacobian_bc<SYMM, OT>), dim3((unsigned)((ns + 255) / 256)), dim3(256),
                     0, st, P.a, J, ns);
}

// RichardsJacobianEval into coefficient arrays of type double (fp32 = 0) or
// float (fp32 = 1); J[7] in stencil order C, W, E, S, N, L, U, with symm
// only C, E, N, U used.
int pfhip_jacobian_into(pfhip_problem *pb, const double *p, double dt, int symm,
                        void *const *J, int fp32, void *stream)
{
  Geom g = geom_of(pb);
  if (pb->overland) {
    int rc = pfhip_overland_flow_eval_kin(pb, p, pb->q2[0], pb->q2[1], pb->q2[2],
                                          pb->q2[3], PFHIP_CALCDER, stream);
    if (rc) return rc;
  }
  JacPFArgs P;
  JacArgs &A = P.a;
  A.g = g;
  A.f = phys_of(pb);
  A.dx = pb->d.dx; A.dy = pb->d.dy; A.dz = pb->d.dz; A.dt = dt;
  A.p = p;
  A.kx = pb->d.kx; A.ky = pb->d.ky; A.kz = pb->d.kz;
  A.poro = pb->d.poro; A.ss = pb->d.ss;
  A.krho = nullptr; A.dkrho = nullptr;
  A.ke = pb->q2[0]; A.kw = pb->q2[1]; A.kn = pb->q2[2]; A.ks = pb->q2[3];
  for (int f = 0; f < 6; f++) { A.bcv[f] = pb->bcv[f]; A.bct[f] = pb->d.bc[f].type; }
  A.symm = symm;
  A.overland = pb->overland;
  for (int s = 0; s < 7; s++) A.J[s] = nullptr;
  P.idx = 1.0 / A.dx; P.idy = 1.0 / A.dy; P.idz = 1.0 / A.dz; P.imu = 1.0 / A.f.mu;
  hipStream_t st = STREAM(stream);
  if (fp32) {
    JOut<float> O;
    for (int s = 0; s < 7; s++) O.c[s] = (float *)J[s];
    if (symm) launch_jacobian_pf<true, float>(P, O, st);
    else launch_jacobian_pf<false, float>(P, O, st);
  } else {
    JOut<double> O;
    for (int s = 0; s < 7; s++) O.c[s] = (double *)J[s];
    if (symm) launch_jacobian_pf<true, double>(P, O, st);
    else launch_jacobian_pf<false, double>(P, O, st);
  }
  LAUNCHCHK();
  return PFHIP_OK;
}

extern "C" int pfhip_richards_jacobian_eval(pfhip_problem *pb, const double *p,
                                            double dt, int symm_part,
                                            double *const *J, void *stream)
{
  if (!J || !J[0] || !J[2] || !J[4] || !J[6] ||
      (!symm_part && (!J[1] || !J[3] || !J[5]))) {
    pfhip_set_error("jacobian_eval: missing coefficient arrays");
    return PFHIP_ERR_ARG;
  }
  void *Jv[7];
  for (int s = 0; s < 7; s++) Jv[s] = J[s];
  return pfhip_jacobian_into(pb, p, dt, symm_part, Jv, 0, stream);
}

// --------------------------------------------------------------------------
// 7-point matvec y = A x (PL/matvec.c:40-330), one pass instead of seven
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_matvec(Geom g, const double *__restrict__ C,
                                                const double *__restrict__ W,
                                                const double *__restrict__ E,
                                                const double *__restrict__ S,
                                                const double *__restrict__ N,
                                                const double *__restrict__ L,
                                                const double *__restrict__ U, int symm,
                                                const double *__restrict__ x,
                                                double *__restrict__ y)
{
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int sy = g.sy, sz = g.sz;
  int ip = (i + 1) + (j + 1) * sy + (k + 1) * sz;
  double v = C[ip] * x[ip];
  if (i > 0) v += (symm ? E[ip - 1] : W[ip]) * x[ip - 1];
  if (i + 1 < g.nx) v += E[ip] * x[ip + 1];
  if (j > 0) v += (symm ? N[ip - sy] : S[ip]) * x[ip - sy];
  if (j + 1 < g.ny) v += N[ip] * x[ip + sy];
  if (k > 0) v += (symm ? U[ip - sz] : L[ip]) * x[ip - sz];
  if (k + 1 < g.nz) v += U[ip] * x[ip + sz];
  y[ip] = v;
}

extern "C" int pfhip_matvec(pfhip_problem *pb, const double *const *A, int symm,
                            const double *x, double *y, void *stream)
{
  Geom g = geom_of(pb);
  hipLaunchKernelGGL(k_matvec, grid3(g), BLOCK3, 0, STREAM(stream), g, A[0],
                     symm ? A[2] : A[1], A[2], symm ? A[4] : A[3], A[4],
                     symm ? A[6] : A[5], A[6], symm, x, y);
  LAUNCHCHK();
  return PFHIP_OK;
}

// ==========================================================================
// Ensemble: many independent problems per launch.  Each member keeps its
// own ghosted arrays; a per-launch table carries every member's geometry
// and pointers.  A member's vector operation (linear combination or
// reduction) runs in one workgroup, so one launch covers every member with a
// pending request and each reduction is finished inside its workgroup in a
// fixed order.  The residual and the Jacobian run as flat grids over all
// members' cells (a block never straddles two members) in two passes: the
// closures of every cell into member scratch, then the face terms of every
// cell from its own and its neighbours' closures.  Members differ in size
// by two orders of magnitude, which the flat grids balance and the
// one-workgroup reductions keep deterministic.
// ==========================================================================

__device__ __forceinline__ int cell_ip(const Geom &g, int c, int &i, int &j, int &k)
{
  i = c % g.nx;
  int r = c / g.nx;
  j = r % g.ny;
  k = r / g.ny;
  return (i + 1) + (j + 1) * g.sy + (k + 1) * g.sz;
}

// the member owning flat block b: last e with starts[e] <= b
__device__ __forceinline__ int ens_find(const int *__restrict__ starts, int n, int b)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// one member's vector operation (PFHIP_EOP_LIN .. PFHIP_EOP_STOP)
struct EnsVec {
  Geom g;
  int op, nq, base, norm, ncell, qoff;   // terms at QC/QX[qoff .. qoff + nq)
  double cst;
  const double *x, *y, *u;
  double *z;
};

// Vector operations as a flat grid: block b takes EVC consecutive interior
// cells of one member (EVB threads, EVC / EVB cells each, a fixed mapping)
// and writes its partial results to part[b * PFHIP_ENS_NRES + v]; the
// finishing kernel then adds (or maxes) each member's block partials in
// block order, so the results are reproducible bit for bit.
#define EVB 256
#define EVC 2048

__device__ __forceinline__ bool ens_is_max(int op)
{
  return op == PFHIP_EOP_MAXABS || op == PFHIP_EOP_SCSTEP || op == PFHIP_EOP_STOP;
}

__device__ __forceinline__ int ens_nres(const EnsVec &E)
{
  switch (E.op) {
    case PFHIP_EOP_LIN: return E.norm ? 1 : 0;
    case PFHIP_EOP_DQ: return 3;
    case PFHIP_EOP_STOP: return 2;
    case PFHIP_EOP_MDOT: return E.nq;
    default: return 1;
  }
}

#define ENS_MDOT 32
__global__ __launch_bounds__(EVB) void k_ens_vec(const EnsVec *__restrict__ T,
                                                 const int *__restrict__ starts, int n,
                                                 const double *__restrict__ QC,
                                                 const double *const *__restrict__ QX,
                                                 double *__restrict__ part)
{
  const int e = ens_find(starts, n, blockIdx.x);
  const EnsVec &E = T[e];
  const double *__restrict__ Ec = QC + E.qoff;
  const double *const *__restrict__ Ex = QX + E.qoff;
  const Geom g = E.g;
  const int op = E.op, nq = E.nq;
  const int c0 = (blockIdx.x - starts[e]) * EVC;
  const int c1 = min(c0 + EVC, E.ncell);
  double *P = part + (long)blockIdx.x * PFHIP_ENS_NRES;
  int i, j, k;
  if (op == PFHIP_EOP_LIN) {
    // z may alias an input: every cell is read before it is written, by
    // the same thread
    double acc = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += EVB) {
      const int ip = cell_ip(g, c, i, j, k);
      double s;
      int q0;
      if (E.base) { s = E.z[ip]; q0 = 0; }
      else if (nq > 0) { s = Ec[0] * Ex[0][ip]; q0 = 1; }
      else { s = E.cst; q0 = 0; }
      for (int q = q0; q < nq; q++) s += Ec[q] * Ex[q][ip];
      E.z[ip] = s;
      acc += s * s;
    }
    if (E.norm) {
      acc = block_reduce<false>(acc);
      if (threadIdx.x == 0) P[0] = acc;
    }
    return;
  }
  if (op == PFHIP_EOP_MDOT) {
    double acc[ENS_MDOT];
#pragma unroll
    for (int q = 0; q < ENS_MDOT; q++) acc[q] = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += EVB) {
      const int ip = cell_ip(g, c, i, j, k);
      const double w = E.y[ip];
#pragma unroll
      for (int q = 0; q < ENS_MDOT; q++)
        if (q < nq) acc[q] += Ex[q][ip] * w;
    }
#pragma unroll
    for (int q = 0; q < ENS_MDOT; q++) {
      if (q < nq) {
        double s = block_reduce<false>(acc[q]);
        if (threadIdx.x == 0) P[q] = s;
      }
    }
    return;
  }
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int c = c0 + threadIdx.x; c < c1; c += EVB) {
    const int ip = cell_ip(g, c, i, j, k);
    switch (op) {
      case PFHIP_EOP_DOT: a0 += E.x[ip] * E.y[ip]; break;
      case PFHIP_EOP_MAXABS: a0 = fmax(a0, fabs(E.x[ip])); break;
      case PFHIP_EOP_SCSTEP: a0 = fmax(a0, fabs(E.y[ip] / (1.0 + fabs(E.x[ip])))); break;
      case PFHIP_EOP_DQ: {
        double u = E.x[ip], v = E.y[ip];
        a0 += u * v;
        a1 += v * v;
        a2 += fabs(v);
        break;
      }
      default: {   // PFHIP_EOP_STOP
        a0 = fmax(a0, fabs(E.x[ip]));
        double un = E.y[ip];
        double st = 1.0 * un + (-1.0) * E.u[ip];
        a1 = fmax(a1, fabs(st / (1.0 + fabs(un))));
      }
    }
  }
  const bool mx = ens_is_max(op);
  a0 = mx ? block_reduce<true>(a0) : block_reduce<false>(a0);
  if (op == PFHIP_EOP_DQ) { a1 = block_reduce<false>(a1); a2 = block_reduce<false>(a2); }
  if (op == PFHIP_EOP_STOP) a1 = block_reduce<true>(a1);
  if (threadIdx.x == 0) { P[0] = a0; P[1] = a1; P[2] = a2; }
}

// one block per request: its member's block partials in block order
__global__ __launch_bounds__(EVB) void k_ens_vec_final(const EnsVec *__restrict__ T,
                                                       const int *__restrict__ starts,
                                                       const double *__restrict__ part,
                                                       double *__restrict__ res)
{
  const int e = blockIdx.x;
  const EnsVec &E = T[e];
  const int nv = ens_nres(E);
  const bool mx = ens_is_max(E.op);
  const int b0 = starts[e], b1 = starts[e + 1];
  for (int v = 0; v < nv; v++) {
    double a = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += EVB) {
      const double t = part[(long)b * PFHIP_ENS_NRES + v];
      a = mx ? fmax(a, t) : a + t;
    }
    a = mx ? block_reduce<true>(a) : block_reduce<false>(a);
    if (threadIdx.x == 0) res[(long)e * PFHIP_ENS_NRES + v] = a;
  }
}

// room for `bytes` of table at the ring's current offset (host address);
// pfhip_ens_commit then copies it to the device twin
extern "C" void *pfhip_ens_reserve(pfhip_ens_stage *st, long bytes, void *stream)
{
  long need = (bytes + 255) & ~255L;
  if (need > st->cap) {
    pfhip_set_error("ensemble table of %ld bytes exceeds the staging ring", bytes);
    return nullptr;
  }
  if (st->off + need > st->cap) {
    if (hipStreamSynchronize(STREAM(stream)) != hipSuccess) {
      pfhip_set_error("ensemble staging: stream synchronisation failed");
      return nullptr;
    }
    st->off = 0;
  }
  return st->host + st->off;
}

extern "C" void *pfhip_ens_commit(pfhip_ens_stage *st, long bytes, void *stream)
{
  long need = (bytes + 255) & ~255L;
  if (hipMemcpyAsync(st->dev + st->off, st->host + st->off, bytes, hipMemcpyHostToDevice,
                     STREAM(stream)) != hipSuccess) {
    pfhip_set_error("ensemble staging: host-to-device copy failed");
    return nullptr;
  }
  void *d = st->dev + st->off;
  st->off += need;
  return d;
}

extern "C" void *pfhip_ens_upload(pfhip_ens_stage *st, const void *src, long bytes,
                                  void *stream)
{
  long need = (bytes + 255) & ~255L;
  if (need > st->cap) {
    pfhip_set_error("ensemble table of %ld bytes exceeds the staging ring", bytes);
    return nullptr;
  }
  if (st->off + need > st->cap) {
    // wrap: earlier copies and the launches reading them must be done
    if (hipStreamSynchronize(STREAM(stream)) != hipSuccess) {
      pfhip_set_error("ensemble staging: stream synchronisation failed");
      return nullptr;
    }
    st->off = 0;
  }
  memcpy(st->host + st->off, src, bytes);
  if (hipMemcpyAsync(st->dev + st->off, st->host + st->off, bytes, hipMemcpyHostToDevice,
                     STREAM(stream)) != hipSuccess) {
    pfhip_set_error("ensemble staging: host-to-device copy failed");
    return nullptr;
  }
  void *d = st->dev + st->off;
  st->off += need;
  return d;
}

static inline long al256(long b) { return (b + 255) & ~255L; }

extern "C" int pfhip_ens_exec_vec(pfhip_ens_req *const *rq, int n, double *res,
                                  double *part, pfhip_ens_stage *st, void *stream)
{
  if (n <= 0) return PFHIP_OK;
  long nqt = 0;
  for (int r = 0; r < n; r++) nqt += rq[r]->nq;
  // one table upload: entries | block starts | term coefficients | term pointers
  const long bT = al256((long)n * sizeof(EnsVec)), bS = al256((long)(n + 1) * sizeof(int));
  const long bC = al256(nqt * (long)sizeof(double));
  const long bytes = bT + bS + bC + al256(nqt * (long)sizeof(void *));
  char *h = (char *)pfhip_ens_reserve(st, bytes, stream);
  if (!h) return PFHIP_ERR_HIP;
  EnsVec *T = (EnsVec *)h;
  int *starts = (int *)(h + bT);
  double *QC = (double *)(h + bT + bS);
  const double **QX = (const double **)(h + bT + bS + bC);
  int results = 0, blk = 0;
  long qo = 0;
  for (int r = 0; r < n; r++) {
    const pfhip_ens_req *q = rq[r];
    EnsVec &E = T[r];
    E.g = geom_of(q->pb);
    E.op = q->op; E.nq = q->nq; E.base = q->base; E.norm = q->norm; E.cst = q->cst;
    E.x = q->x; E.y = q->y; E.u = q->u; E.z = q->z;
    E.ncell = (int)q->pb->ncell;
    E.qoff = (int)qo;
    for (int t = 0; t < q->nq; t++) { QC[qo + t] = q->c[t]; QX[qo + t] = q->xq[t]; }
    qo += q->nq;
    starts[r] = blk;
    blk += (int)((q->pb->ncell + EVC - 1) / EVC);
    if (q->nout) results = 1;
  }
  starts[n] = blk;
  char *d = (char *)pfhip_ens_commit(st, bytes, stream);
  if (!d) return PFHIP_ERR_HIP;
  const EnsVec *dT = (const EnsVec *)d;
  const int *dS = (const int *)(d + bT);
  hipLaunchKernelGGL(k_ens_vec, dim3(blk), dim3(EVB), 0, STREAM(stream), dT, dS, n,
                     (const double *)(d + bT + bS), (const double *const *)(d + bT + bS + bC),
                     part);
  if (results)
    hipLaunchKernelGGL(k_ens_vec_final, dim3(n), dim3(EVB), 0, STREAM(stream), dT, dS,
                       (const double *)part, res);
  LAUNCHCHK();
  return PFHIP_OK;
}

// blocks of the flat vector grid of a member of ncell cells
extern "C" long pfhip_ens_vec_blocks(long ncell) { return (ncell + EVC - 1) / EVC; }

// ---- flat two-pass residual ------------------------------------------------
#define ERB 256   // cells per block of the flat member grids

struct EnsRes {
  ResArgs A;
  double *pe, *kr, *sat;   // member scratch: evaluation pressure, kr rho, S
  int mode, ncell;
};

// pass 1: pressure of the form (p, or p + sig pv), S and kr rho of every cell
__global__ __launch_bounds__(ERB) void k_ens_res_closure(const EnsRes *__restrict__ T,
                                                         const int *__restrict__ starts,
                                                         int n)
{
  const int e = ens_find(starts, n, blockIdx.x);
  const EnsRes &E = T[e];
  const int c = (blockIdx.x - starts[e]) * ERB + threadIdx.x;
  if (c >= E.ncell) return;
  const ResArgs &A = E.A;
  const Phys &f = A.f;
  int i, j, k;
  const int ip = cell_ip(A.g, c, i, j, k);
  const double p = (E.mode != 0) ? A.p[ip] + A.psig * A.pv[ip] : A.p[ip];
  double S, kr;
  vg_s_kr(p, f.irg, f.alpha[ip], f.nvg[ip], f.Sres(ip), f.Ssat(ip), S, kr, f.tb);
  E.pe[ip] = p;
  E.kr[ip] = kr * f.rho;
  E.sat[ip] = S;
  if (E.mode == 2) A.pout[ip] = p;
}

// pass 2: the cell's six faces (own upper faces and the lower faces, each
// with the lower cell first, as the tiled kernel evaluates them), storage,
// the boundary-face terms, and the output form
__global__ __launch_bounds__(ERB) void k_ens_res_faces(const EnsRes *__restrict__ T,
                                                       const int *__restrict__ starts,
                                                       int n)
{
  const int e = ens_find(starts, n, blockIdx.x);
  const EnsRes &E = T[e];
  const int c = (blockIdx.x - starts[e]) * ERB + threadIdx.x;
  if (c >= E.ncell) return;
  const ResArgs &A = E.A;
  const Geom &g = A.g;
  const Phys &f = A.f;
  int i, j, k;
  const int ip = cell_ip(g, c, i, j, k);
  const int sy = g.sy, sz = g.sz;
  const double grav = f.g, imu = A.imu;
  const double hrg = 0.5 * f.rho * grav * 1.0;
  const double ffx = A.dy * A.dz, ffy = A.dx * A.dz, ffz = A.dx * A.dy;
  const double *__restrict__ P = E.pe;
  const double *__restrict__ KR = E.kr;
  const double pc = P[ip], krc = KR[ip];
  const double kxc = A.kx[ip], kyc = A.ky[ip], kzc = A.kz[ip];
  double ur = 0.0, uf = 0.0, uu = 0.0, ul = 0.0, ud = 0.0, ub = 0.0;
  if (i + 1 < g.nx)
    ur = flux_xy_i(pc, P[ip + 1], kxc, A.kx[ip + 1], krc, KR[ip + 1], ffx, A.idx, grav, imu);
  if (j + 1 < g.ny)
    uf = flux_xy_i(pc, P[ip + sy], kyc, A.ky[ip + sy], krc, KR[ip + sy], ffy, A.idy, grav, imu);
  if (k + 1 < g.nz)
    uu = flux_z_i(pc, P[ip + sz], kzc, A.kz[ip + sz], krc, KR[ip + sz], ffz, A.idz, hrg, imu);
  if (i > 0)
    ul = flux_xy_i(P[ip - 1], pc, A.kx[ip - 1], kxc, KR[ip - 1], krc, ffx, A.idx, grav, imu);
  if (j > 0)
    ud = flux_xy_i(P[ip - sy], pc, A.ky[ip - sy], kyc, KR[ip - sy], krc, ffy, A.idy, grav, imu);
  if (k > 0)
    ub = flux_z_i(P[ip - sz], pc, A.kz[ip - sz], kzc, KR[ip - sz], krc, ffz, A.idz, hrg, imu);
  double fv = residual_cell(A, i, j, k, ip, pc, krc, E.sat[ip], kxc, kyc, kzc, ur, uf, uu, ul,
                            ud, ub, A.s_old[ip], A.poro[ip], A.ss[ip], A.p_old[ip]);
  const bool bnd = (i == 0) | (i == g.nx - 1) | (j == 0) | (j == g.ny - 1) | (k == 0) |
                   (k == g.nz - 1);
  if (bnd) {
    const double krb = vg_kr(pc, f.irg, f.alpha[ip], f.nvg[ip], f.tb) * f.rho;
    fv = residual_bc(A, i, j, k, ip, fv, pc, krb, kxc, kyc, kzc);
  }
  if (E.mode == 1) fv = A.fsc * fv - A.fsc * A.fsub[ip];
  A.F[ip] = fv;
}

extern "C" int pfhip_ens_exec_residual(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                                       void *stream)
{
  if (n <= 0) return PFHIP_OK;
  const long bT = al256((long)n * sizeof(EnsRes));
  const long bytes = bT + al256((long)(n + 1) * sizeof(int));
  char *h = (char *)pfhip_ens_reserve(st, bytes, stream);
  if (!h) return PFHIP_ERR_HIP;
  EnsRes *T = (EnsRes *)h;
  int *starts = (int *)(h + bT);
  int blk = 0;
  for (int r = 0; r < n; r++) {
    const pfhip_ens_req *q = rq[r];
    EnsRes &E = T[r];
    E.A = res_args(q->pb, q->xq[0], q->xq[1], q->dt);
    E.A.p = q->x; E.A.F = q->z;
    if (q->mode == 1) { E.A.pv = q->y; E.A.fsub = q->u; E.A.psig = q->sig; E.A.fsc = q->fsc; }
    if (q->mode == 2) { E.A.pv = q->y; E.A.pout = q->z2; E.A.psig = q->sig; }
    E.pe = q->scratch[0]; E.kr = q->scratch[1]; E.sat = q->scratch[2];
    E.mode = q->mode;
    E.ncell = (int)q->pb->ncell;
    starts[r] = blk;
    blk += (int)((q->pb->ncell + ERB - 1) / ERB);
  }
  starts[n] = blk;
  char *d = (char *)pfhip_ens_commit(st, bytes, stream);
  if (!d) return PFHIP_ERR_HIP;
  hipLaunchKernelGGL(k_ens_res_closure, dim3(blk), dim3(ERB), 0, STREAM(stream),
                     (const EnsRes *)d, (const int *)(d + bT), n);
  hipLaunchKernelGGL(k_ens_res_faces, dim3(blk), dim3(ERB), 0, STREAM(stream),
                     (const EnsRes *)d, (const int *)(d + bT), n);
  LAUNCHCHK();
  return PFHIP_OK;
}

// ---- flat two-pass PFSymmetric Jacobian into the members' level-0 matrices --
struct EnsJac {
  JacArgs A;
  double *kr, *dk, *sat, *dsat;   // member scratch
  void *J[4];                     // C, E, N, U of the multigrid's finest level
  int ncell;
};

__global__ __launch_bounds__(ERB) void k_ens_jac_closure(const EnsJac *__restrict__ T,
                                                         const int *__restrict__ starts,
                                                         int n)
{
  const int e = ens_find(starts, n, blockIdx.x);
  const EnsJac &E = T[e];
  const int c = (blockIdx.x - starts[e]) * ERB + threadIdx.x;
  if (c >= E.ncell) return;
  const JacArgs &A = E.A;
  const Phys &f = A.f;
  int i, j, k;
  const int ip = cell_ip(A.g, c, i, j, k);
  double S, dS, kr, dk;
  vg_all(A.p[ip], f.irg, f.alpha[ip], f.nvg[ip], f.Sres(ip), f.Ssat(ip), S, dS, kr, dk, f.tb);
  kr *= f.rho;
  E.kr[ip] = kr;
  E.dk[ip] = dk * f.rho + kr * 0.0;
  E.sat[ip] = S;
  E.dsat[ip] = dS;
}

template <typename OT>
__global__ __launch_bounds__(ERB) void k_ens_jac_rows(const EnsJac *__restrict__ T,
                                                      const int *__restrict__ starts, int n)
{
  const int e = ens_find(starts, n, blockIdx.x);
  const EnsJac &E = T[e];
  const int c = (blockIdx.x - starts[e]) * ERB + threadIdx.x;
  if (c >= E.ncell) return;
  const JacArgs &A = E.A;
  const Geom &g = A.g;
  const Phys &f = A.f;
  int i, j, k;
  const int ip = cell_ip(g, c, i, j, k);
  const int sy = g.sy, sz = g.sz;
  const double rho = f.rho, gc = f.g, imu = 1.0 / f.mu;
  const double dt = A.dt, dx = A.dx, dy = A.dy, dz = A.dz;
  const double idx = 1.0 / dx, idy = 1.0 / dy;
  const double isep = 1.0 / (dz * 0.5 * (1.0 + 1.0));
  const double hrg = (1.0 / (1.0 + 1.0)) * rho * gc;
  const double vol2 = dx * dy * dz * 1.0;
  const double ffx = dy * dz, ffy = dx * dz, ffz = dx * dy;
  const double ddp = 0.0;
  const double *__restrict__ P = A.p;
  const double *__restrict__ KR = E.kr;
  const double *__restrict__ DK = E.dk;
  const double pc = P[ip], krc = KR[ip], dkc = DK[ip], S = E.sat[ip], dS = E.dsat[ip];
  double cc = (dS * rho + S * ddp) * A.poro[ip] * vol2 +
              A.ss[ip] * vol2 * (dS * rho * pc + S * ddp * pc + S * rho);
  double offE = 0.0, offN = 0.0, offU = 0.0;
  double T1, T2, sym, syml, symu;
  if (i + 1 < g.nx) {
    jface_xy(pc, P[ip + 1], A.kx[ip], A.kx[ip + 1], krc, KR[ip + 1], dkc, DK[ip + 1], dt, ffx,
             idx, gc, imu, T1, T2, sym);
    cc -= T1;
    offE = sym;
  }
  if (i > 0) {
    jface_xy(P[ip - 1], pc, A.kx[ip - 1], A.kx[ip], KR[ip - 1], krc, DK[ip - 1], dkc, dt, ffx,
             idx, gc, imu, T1, T2, sym);
    cc -= T2;
  }
  if (j + 1 < g.ny) {
    jface_xy(pc, P[ip + sy], A.ky[ip], A.ky[ip + sy], krc, KR[ip + sy], dkc, DK[ip + sy], dt,
             ffy, idy, gc, imu, T1, T2, sym);
    cc -= T1;
    offN = sym;
  }
  if (j > 0) {
    jface_xy(P[ip - sy], pc, A.ky[ip - sy], A.ky[ip], KR[ip - sy], krc, DK[ip - sy], dkc, dt,
             ffy, idy, gc, imu, T1, T2, sym);
    cc -= T2;
  }
  if (k + 1 < g.nz) {
    jface_z(pc, P[ip + sz], A.kz[ip], A.kz[ip + sz], krc, KR[ip + sz], dkc, DK[ip + sz], dt,
            ffz, isep, hrg, imu, T1, T2, syml, symu);
    cc -= T1;
    offU = symu;
  }
  if (k > 0) {
    jface_z(P[ip - sz], pc, A.kz[ip - sz], A.kz[ip], KR[ip - sz], krc, DK[ip - sz], dkc, dt,
            ffz, isep, hrg, imu, T1, T2, syml, symu);
    cc -= T2;
  }
  const bool dir = (i == 0 && A.bct[0] == PFHIP_BC_DIR_EQUIL) ||
                   (i == g.nx - 1 && A.bct[1] == PFHIP_BC_DIR_EQUIL) ||
                   (j == 0 && A.bct[2] == PFHIP_BC_DIR_EQUIL) ||
                   (j == g.ny - 1 && A.bct[3] == PFHIP_BC_DIR_EQUIL) ||
                   (k == 0 && A.bct[4] == PFHIP_BC_DIR_EQUIL) ||
                   (k == g.nz - 1 && A.bct[5] == PFHIP_BC_DIR_EQUIL);
  // the single-problem path stores the interior row, then applies the
  // Dirichlet faces to the stored (rounded) diagonal
  if (dir) cc = jac_dir_faces(A, i, j, k, ip, (double)(OT)cc);
  ((OT *)E.J[0])[ip] = (OT)cc;
  ((OT *)E.J[1])[ip] = (OT)offE;
  ((OT *)E.J[2])[ip] = (OT)offN;
  ((OT *)E.J[3])[ip] = (OT)offU;
}

extern "C" int pfhip_ens_exec_jacobian(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                                       void *stream)
{
  if (n <= 0) return PFHIP_OK;
  const long bT = al256((long)n * sizeof(EnsJac));
  const long bytes = bT + al256((long)(n + 1) * sizeof(int));
  char *h = (char *)pfhip_ens_reserve(st, bytes, stream);
  if (!h) return PFHIP_ERR_HIP;
  EnsJac *T = (EnsJac *)h;
  int *starts = (int *)(h + bT);
  int prec0 = 0, blk = 0;
  for (int r = 0; r < n; r++) {
    const pfhip_ens_req *q = rq[r];
    pfhip_problem *pb = q->pb;
    EnsJac &E = T[r];
    JacArgs &A = E.A;
    A.g = geom_of(pb);
    A.f = phys_of(pb);
    A.dx = pb->d.dx; A.dy = pb->d.dy; A.dz = pb->d.dz; A.dt = q->dt;
    A.p = q->x;
    A.kx = pb->d.kx; A.ky = pb->d.ky; A.kz = pb->d.kz;
    A.poro = pb->d.poro; A.ss = pb->d.ss;
    A.krho = nullptr; A.dkrho = nullptr;
    A.ke = A.kw = A.kn = A.ks = nullptr;
    for (int f = 0; f < 6; f++) { A.bcv[f] = pb->bcv[f]; A.bct[f] = pb->d.bc[f].type; }
    A.symm = 1;
    A.overland = 0;
    for (int k = 0; k < 7; k++) A.J[k] = nullptr;
    E.kr = q->scratch[0]; E.dk = q->scratch[1]; E.sat = q->scratch[2]; E.dsat = q->scratch[3];
    int prec;
    pfhip_mg_level0(q->mg, E.J, &prec);
    if (r == 0) prec0 = prec;
    if (prec != prec0 || pb->overland) {
      pfhip_set_error("ensemble members need one multigrid precision and no overland flow");
      return PFHIP_ERR_ARG;
    }
    E.ncell = (int)pb->ncell;
    starts[r] = blk;
    blk += (int)((pb->ncell + ERB - 1) / ERB);
  }
  starts[n] = blk;
  char *d = (char *)pfhip_ens_commit(st, bytes, stream);
  if (!d) return PFHIP_ERR_HIP;
  const EnsJac *dT = (const EnsJac *)d;
  const int *dS = (const int *)(d + bT);
  hipLaunchKernelGGL(k_ens_jac_closure, dim3(blk), dim3(ERB), 0, STREAM(stream), dT, dS, n);
  if (prec0 == 32)
    hipLaunchKernelGGL(k_ens_jac_rows<float>, dim3(blk), dim3(ERB), 0, STREAM(stream), dT, dS, n);
  else
    hipLaunchKernelGGL(k_ens_jac_rows<double>, dim3(blk), dim3(ERB), 0, STREAM(stream), dT, dS,
                       n);
  LAUNCHCHK();
  return PFHIP_OK;
}

// pinned staging ring, device result slots
extern "C" int pfhip_ens_stage_create(pfhip_ens_stage *st, long cap)
{
  st->off = 0;
  st->cap = cap;
  st->host = nullptr;
  st->dev = nullptr;
  HIPCHK(hipHostMalloc((void **)&st->host, cap, hipHostMallocDefault));
  HIPCHK(hipMalloc((void **)&st->dev, cap));
  return PFHIP_OK;
}

extern "C" void pfhip_ens_stage_destroy(pfhip_ens_stage *st)
{
  if (st->host) hipHostFree(st->host);
  if (st->dev) hipFree(st->dev);
  st->host = st->dev = nullptr;
}

extern "C" int pfhip_ens_fetch(double *host, const double *dev, long n, void *stream)
{
  if (n > 0)
    HIPCHK(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost,
                          STREAM(stream)));
  HIPCHK(hipStreamSynchronize(STREAM(stream)));
  return PFHIP_OK;
}

extern "C" int pfhip_host_alloc(double **p, long n)
{
  *p = nullptr;
  HIPCHK(hipHostMalloc((void **)p, sizeof(double) * (n > 0 ? n : 1), hipHostMallocDefault));
  return PFHIP_OK;
}

extern "C" void pfhip_host_free(double *p)
{
  if (p) hipHostFree(p);
}

// ---- BCPressure and Saturation of many members -----------------------------
// BC values of every requesting member are computed on the host into one
// staged block and scattered to the members' face tables by one kernel
// (one segment of one face per block).
struct EnsSeg {
  double *dst;
  long src;      // offset in the staged values
  int n;
};

__global__ __launch_bounds__(256) void k_ens_scatter(const EnsSeg *__restrict__ S,
                                                     const double *__restrict__ vals)
{
  const EnsSeg &G = S[blockIdx.x];
  for (int c = threadIdx.x; c < G.n; c += blockDim.x) G.dst[c] = vals[G.src + c];
}

extern "C" int pfhip_ens_exec_bc(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                                 void *stream)
{
  if (n <= 0) return PFHIP_OK;
  long nvals = 0;
  for (int r = 0; r < n; r++)
    for (int f = 0; f < 6; f++) nvals += face_ncells(&rq[r]->pb->d, f);
  const long bS = al256((long)6 * n * sizeof(EnsSeg));
  const long bytes = bS + al256(nvals * (long)sizeof(double));
  char *h = (char *)pfhip_ens_reserve(st, bytes, stream);
  if (!h) return PFHIP_ERR_HIP;
  EnsSeg *S = (EnsSeg *)h;
  double *vals = (double *)(h + bS);
  long off = 0;
  for (int r = 0; r < n; r++) {
    pfhip_problem *pb = rq[r]->pb;
    for (int f = 0; f < 6; f++) {
      EnsSeg &G = S[6 * r + f];
      G.n = face_ncells(&pb->d, f);
      G.dst = pb->bcv[f];
      G.src = off;
      bc_face_values(pb, f, rq[r]->dt, vals + off);
      off += G.n;
    }
  }
  char *d = (char *)pfhip_ens_commit(st, bytes, stream);
  if (!d) return PFHIP_ERR_HIP;
  hipLaunchKernelGGL(k_ens_scatter, dim3(6 * n), dim3(256), 0, STREAM(stream),
                     (const EnsSeg *)d, (const double *)(d + bS));
  LAUNCHCHK();
  return PFHIP_OK;
}

struct EnsSat {
  Geom g;
  Phys f;
  const double *p;
  double *out;
  int what, ncell;   // what: 0 S, 1 dS/dp (the k_closure codes)
};

__global__ __launch_bounds__(ERB) void k_ens_sat(const EnsSat *__restrict__ T,
                                                 const int *__restrict__ starts, int n)
{
  const int e = ens_find(starts, n, blockIdx.x);
  const EnsSat &E = T[e];
  const int c = (blockIdx.x - starts[e]) * ERB + threadIdx.x;
  if (c >= E.ncell) return;
  const Phys &f = E.f;
  int i, j, k;
  const int ip = cell_ip(E.g, c, i, j, k);
  const double pv = E.p[ip], a = f.alpha[ip], nn = f.nvg[ip];
  E.out[ip] = (E.what == 1) ? vg_dsat(pv, f.irg, a, nn, f.Sres(ip), f.Ssat(ip))
                            : vg_sat(pv, f.irg, a, nn, f.Sres(ip), f.Ssat(ip));
}

extern "C" int pfhip_ens_exec_saturation(pfhip_ens_req *const *rq, int n, pfhip_ens_stage *st,
                                         void *stream)
{
  if (n <= 0) return PFHIP_OK;
  const long bT = al256((long)n * sizeof(EnsSat));
  const long bytes = bT + al256((long)(n + 1) * sizeof(int));
  char *h = (char *)pfhip_ens_reserve(st, bytes, stream);
  if (!h) return PFHIP_ERR_HIP;
  EnsSat *T = (EnsSat *)h;
  int *starts = (int *)(h + bT);
  int blk = 0;
  for (int r = 0; r < n; r++) {
    const pfhip_ens_req *q = rq[r];
    EnsSat &E = T[r];
    E.g = geom_of(q->pb);
    E.f = phys_of(q->pb);
    E.p = q->x;
    E.out = q->z;
    E.what = q->mode == PFHIP_CALCDER ? 1 : 0;
    E.ncell = (int)q->pb->ncell;
    starts[r] = blk;
    blk += (int)((q->pb->ncell + ERB - 1) / ERB);
  }
  starts[n] = blk;
  char *d = (char *)pfhip_ens_commit(st, bytes, stream);
  if (!d) return PFHIP_ERR_HIP;
  hipLaunchKernelGGL(k_ens_sat, dim3(blk), dim3(ERB), 0, STREAM(stream), (const EnsSat *)d,
                     (const int *)(d + bT), n);
  LAUNCHCHK();
  return PFHIP_OK;
}
