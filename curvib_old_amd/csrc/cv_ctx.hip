// cv_ctx.hip -- context life cycle, device memory, field exchange and the
// deterministic reductions of the CURVIB fractional-step library (gfx950).
#include "cv_internal.h"
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <mutex>

static thread_local char g_err[512];

void cv_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" int cv_last_error(char *buf, int len) {
  if (!buf || len <= 0) return CV_ERR_ARG;
  strncpy(buf, g_err, (size_t)len - 1);
  buf[len - 1] = 0;
  return CV_OK;
}

int cv_alloc(cv_ctx *c, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    cv_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    (void)hipGetLastError();
    return CV_ERR_OOM;
  }
  c->allocs.push_back(*p);
  c->bytes_alloc += bytes;
  return CV_OK;
}

void cv_free(cv_ctx *c, void *p) {
  for (size_t q = 0; q < c->allocs.size(); q++)
    if (c->allocs[q] == p) {
      c->allocs.erase(c->allocs.begin() + q);
      break;
    }
  size_t sz = 0;
  if (hipMemPtrGetInfo(p, &sz) == hipSuccess) c->bytes_alloc -= sz;
  (void)hipFree(p);
}

__global__ void k_copy(double *dst, const double *src, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    dst[t] = src[t];
}
__global__ void k_fill(double *dst, double v, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    dst[t] = v;
}
__global__ void k_axpby(double *y, double a, const double *x, double b, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    y[t] = a * x[t] + b * y[t];
}

extern "C" void cv_default_solver(cv_solver_opts *o) {
  o->precision = CV_FP64;
  o->restart = 30;
  o->max_it = 200;
  o->rtol = 1e-10;
  o->atol = 0.0;
  o->pre_sweeps = 2;
  o->post_sweeps = 2;
  o->coarse_sweeps = 40;
  o->smoother = CV_SMOOTH_JACOBI;
  o->omega = 0.8;
}

extern "C" long cv_device_bytes(const cv_ctx *c) { return c ? (long)c->bytes_alloc : 0; }

extern "C" void cv_default_momentum(cv_momentum_opts *o) {
  o->scheme = CV_MOM_RK;
  o->imp_max_it = 50;
  o->imp_atol = 1e-7;
  o->imp_rtol = 1e-4;
  o->cfl = 0.2;
}

extern "C" int cv_set_momentum(cv_ctx *c, const cv_momentum_opts *o) {
  if (!c || !o || (o->scheme != CV_MOM_RK && o->scheme != CV_MOM_IMPRK) || o->imp_max_it < 1 || o->cfl <= 0)
    return CV_ERR_ARG;
  c->mom = *o;
  return CV_OK;
}

// the analytic inflow profiles of InflowFlux (bcs.c:158-196); 0, 3, 5, 6
// and 8 interpolate an inlet file and are not supported
static bool inlet_profile_ok(int p) {
  return p == 1 || p == -1 || p == 2 || p == 4 || p == 7 || p == 10 || p == 11;
}

// what is wrong with a grid description, or nullptr
static const char *desc_problem(const cv_grid_desc *d) {
  if (!d) return "no description";
  if (d->IM < 4 || d->JM < 4 || d->KM < 4) return "fewer than 4 nodes on an axis";
  for (int q = 0; q < 6; q++) {
    const int b = d->bctype[q];
    if (b == 1 || b == 3 || b == 6 || b == 7) continue;
    if ((q == 4 && b == 5) || (q == 5 && b == 4)) continue;
    return "faces take bctype 1, 3, 6 or 7; the k- face also 5, the k+ face also 4";
  }
  if (d->bctype[2] == 7 || d->bctype[3] == 7) return "j-periodic grids are not supported";
  // FormMetrics leaves eta of the j = 0 face at 0 (its loop starts at j = 1,
  // metrics.c:107-110): slip and far field divide by it
  if (d->bctype[2] == 3 || d->bctype[2] == 6) return "slip and far field are not defined on the j- face";
  if ((d->bctype[0] == 7) != (d->bctype[1] == 7)) return "bctype 7 must pair on the i faces";
  if ((d->bctype[4] == 7) != (d->bctype[5] == 7)) return "bctype 7 must pair on the k faces";
  if (d->bctype[4] == 5 && !inlet_profile_ok(d->inletprofile))
    return "inlet profile not supported (analytic profiles 1, -1, 2, 4, 7, 10, 11)";
  if (d->mglevels < 1 || d->mglevels > CV_MAXLEV) return "mglevels out of range";
  if (d->dt <= 0 || d->ren <= 0) return "dt and ren must be positive";
  if ((long)(d->IM + 1) * (d->JM + 1) * (d->KM + 1) >= (1L << 31)) return "grid too large";   // 32-bit storage indices in kernels
  return nullptr;
}
static int check_desc(const cv_grid_desc *d) { return desc_problem(d) == nullptr; }
// a face that needs Ubcs storage: inlet, outlet, slip or far field
static bool desc_ubf(const cv_grid_desc *d) {
  for (int q = 0; q < 6; q++)
    if (d->bctype[q] != 1 && d->bctype[q] != 7) return true;
  return false;
}

// MGDACreate dimensions (init.c:78-99), level 0 coarsest; each coarsening
// must keep an even cell count on coarsened axes so that fine cells pair
// up.  With coarsen_auto a transition coarsens, among the axes the semi
// flags allow, only those with the most cells; each level stores the flags
// of the transition that produced it.  Returns 0 or -1 with the error text
// set.
static int level_dims(const cv_grid_desc *desc, CvDims *dims) {
  int im = desc->IM, jm = desc->JM, km = desc->KM;
  const bool autoc = desc->coarsen_auto != 0;
  for (int l = desc->mglevels - 1; l >= 0; l--) {
    int isc = desc->isc, jsc = desc->jsc, ksc = desc->ksc;
    if (autoc && l < desc->mglevels - 1) {
      // coarsen, among the axes allowed to coarsen, only those with the most cells
      int mxc = 0;
      if (!isc) mxc = std::max(mxc, im - 1);
      if (!jsc) mxc = std::max(mxc, jm - 1);
      if (!ksc) mxc = std::max(mxc, km - 1);
      if (!isc && im - 1 < mxc) isc = 1;
      if (!jsc && jm - 1 < mxc) jsc = 1;
      if (!ksc && km - 1 < mxc) ksc = 1;
    }
    if (l < desc->mglevels - 1) {
      if ((!isc && (im - 1) % 2) || (!jsc && (jm - 1) % 2) || (!ksc && (km - 1) % 2)) {
        cv_set_error("cv_ctx_create: grid cannot be restricted to %d levels", desc->mglevels);
        return -1;
      }
      im = isc ? im : (im + 1) / 2;
      jm = jsc ? jm : (jm + 1) / 2;
      km = ksc ? km : (km + 1) / 2;
    }
    if (im < 4 || jm < 4 || km < 4) {
      cv_set_error("cv_ctx_create: level %d too small (%d %d %d nodes)", l, im, jm, km);
      return -1;
    }
    CvDims &D = dims[l];
    D.mx = im + 1; D.my = jm + 1; D.mz = km + 1;
    D.n = (long)D.mx * D.my * D.mz;
    D.xper = desc->bctype[0] == 7;
    D.zper = desc->bctype[4] == 7;
    D.isc = isc; D.jsc = jsc; D.ksc = ksc;
  }
  return 0;
}

// device bytes of one context: per level (FP64, 22 arrays: nvert, aj, 9 g, dinv, 3
// coordinates, 3 line factors, line scratch, x b r; FP32: 9 g, aj, dinv, 3
// line factors, line scratch, x b r (18 arrays); the operator word), the finest-level
// metrics (40 doubles), state (25 doubles: 7 vectors of 3, p phi b w) and
// the 3n exchange staging, reduction scratch and the Krylov space of
// FGMRES(restart); with an inlet, outlet, slip or far-field face the Ubcs
// of the six boundary planes
static size_t plan_bytes(const CvDims *dims, int nlev, int restart, bool ubf) {
  size_t b = 0;
  const CvDims &F = dims[nlev - 1];
  if (ubf) b += (size_t)3 * 2 * ((size_t)F.my * F.mz + (size_t)F.mx * F.mz + (size_t)F.mx * F.my) * 8;
  for (int l = 0; l < nlev; l++) b += (size_t)dims[l].n * (22 * 8 + 19 * 4 + 4);
  const size_t n = (size_t)dims[nlev - 1].n;
  b += n * 8 * (40 + 25 + 3);
  b += (size_t)(1 << 16) * 8 + 128 * 8 + 34 * 1024 * 8;
  if (restart < 1) restart = 1;
  if (restart > 31) restart = 31;
  b += n * 8 * (size_t)(2 * restart + 1);
  return b;
}

// device bytes of the CV_MIXED16 FP16 packs (32 per point and level)
long cv_half_bytes(const cv_grid_desc *desc) {
  CvDims dims[CV_MAXLEV];
  if (!check_desc(desc) || level_dims(desc, dims)) return 0;
  long b = 0;
  for (int l = 0; l < desc->mglevels; l++) b += dims[l].n * 32;
  return b;
}

extern "C" long cv_ctx_bytes(const cv_grid_desc *desc, int restart) {
  if (!check_desc(desc)) return -1;
  CvDims dims[CV_MAXLEV];
  if (level_dims(desc, dims)) return -1;
  return (long)plan_bytes(dims, desc->mglevels, restart, desc_ubf(desc));
}

extern "C" int cv_ctx_create(const cv_grid_desc *desc, int device, cv_ctx **out) {
  if (!out) return CV_ERR_ARG;
  *out = nullptr;
  if (const char *why = desc_problem(desc)) {
    cv_set_error("cv_ctx_create: unsupported grid description: %s", why);
    return CV_ERR_ARG;
  }
  CvDims dims[CV_MAXLEV];
  if (level_dims(desc, dims)) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(device));
  cv_ctx *c = new cv_ctx();
  c->device = device;
  c->nlev = desc->mglevels;
  c->IM = desc->IM; c->JM = desc->JM; c->KM = desc->KM;
  memcpy(c->bctype, desc->bctype, sizeof(c->bctype));
  c->inlet = desc->bctype[4] == 5;
  c->outlet = desc->bctype[5] == 4;
  c->inletprofile = desc->inletprofile;
  for (int q = 0; q < 3; q++) c->cm_c[q] = desc->inlet_center[q];
  c->slip = c->far = 0;
  for (int q = 0; q < 6; q++) {
    if (desc->bctype[q] == 3) c->slip |= 1 << q;
    if (desc->bctype[q] == 6) c->far |= 1 << q;
  }
  c->ubf = nullptr;
  c->ren = desc->ren; c->dt = desc->dt; c->st = desc->st > 0 ? desc->st : 1.0;
  cv_default_solver(&c->opt);
  cv_default_momentum(&c->mom);
  for (int l = 0; l < c->nlev; l++) c->lev[l].d = dims[l];
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    cv_set_error("cv_ctx_create: stream creation failed");
    return CV_ERR_HIP;
  }
  int rc;
  // a failing HIP call inside the set-up frees what was allocated so far
#define CV_TRY(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      cv_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));           \
      rc = CV_ERR_HIP;                                                                            \
      goto fail;                                                                                  \
    }                                                                                             \
  } while (0)
  for (int q = 0; q < 4; q++) CV_TRY(hipEventCreate(&c->step_ev[q]));
  // per-level operator storage
  for (int l = 0; l < c->nlev; l++) {
    CvLevel &L = c->lev[l];
    long n = L.d.n;
    void *p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.nvert = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.aj = (double *)p;
    for (int q = 0; q < 9; q++) { if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.g[q] = (double *)p; }
    if ((rc = cv_alloc(c, &p, n * sizeof(uint32_t)))) goto fail; L.flags = (uint32_t *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.dinv = (double *)p;
    for (int q = 0; q < 9; q++) { if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.gf[q] = (float *)p; }
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.ajf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.dinvf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.nwf = (float *)p;
    if ((rc = cv_alloc(c, &p, 3 * n * sizeof(double)))) goto fail; L.coor = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.ylo = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.yip = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.ycp = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.ylof = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.yipf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.ycpf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.ltmp = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.ltmpf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.x = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.b = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; L.r = (double *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.xf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.bf = (float *)p;
    if ((rc = cv_alloc(c, &p, n * sizeof(float)))) goto fail; L.rf = (float *)p;
    CV_TRY(hipMemsetAsync(L.nvert, 0, n * sizeof(double), c->stream));
  }
  {
    const long n = c->lev[c->nlev - 1].d.n;
    void *p;
    CvMetrics &M = c->m;
    double **vec3[] = {M.csi, M.eta, M.zet, M.icsi, M.ieta, M.izet, M.jcsi, M.jeta, M.jzet,
                       M.kcsi, M.keta, M.kzet};
    for (auto v : vec3)
      for (int q = 0; q < 3; q++) { if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; v[q] = (double *)p; }
    double **sc[] = {&M.aj, &M.iaj, &M.jaj, &M.kaj};
    for (auto s : sc) { if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail; *s = (double *)p; }
    M.aj = c->lev[c->nlev - 1].aj ? M.aj : M.aj;
    double **st3[] = {c->ucont, c->ucont_o, c->ucat, c->rhs, c->rct, c->ucont_rm1, c->pucont};
    for (auto v : st3)
      for (int q = 0; q < 3; q++) {
        if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail;
        v[q] = (double *)p;
        CV_TRY(hipMemsetAsync(v[q], 0, n * sizeof(double), c->stream));
      }
    if ((rc = cv_alloc(c, &p, 3 * n * sizeof(double)))) goto fail;
    c->stage = (double *)p;
    double **st1[] = {&c->p, &c->phi, &c->bvec, &c->w};
    for (auto s : st1) {
      if ((rc = cv_alloc(c, &p, n * sizeof(double)))) goto fail;
      *s = (double *)p;
      CV_TRY(hipMemsetAsync(*s, 0, n * sizeof(double), c->stream));
    }
    c->red_len = 1 << 16;
    if ((rc = cv_alloc(c, &p, c->red_len * sizeof(double)))) goto fail; c->red = (double *)p;
    if ((rc = cv_alloc(c, &p, 128 * sizeof(double)))) goto fail; c->dsc = (double *)p;
    CV_TRY(hipMemsetAsync(c->dsc, 0, 128 * sizeof(double), c->stream));   // FluxInSum starts at 0
    if (desc_ubf(desc)) {
      const CvDims &D = c->lev[c->nlev - 1].d;
      const size_t nb = (size_t)6 * ((size_t)D.my * D.mz + (size_t)D.mx * D.mz + (size_t)D.mx * D.my) * sizeof(double);
      if ((rc = cv_alloc(c, &p, nb))) goto fail;
      c->ubf = (double *)p;
      CV_TRY(hipMemsetAsync(c->ubf, 0, nb, c->stream));
    }
    if ((rc = cv_alloc(c, &p, 34 * 1024 * sizeof(double)))) goto fail; c->part = (double *)p;
    c->krylov_m = 0;
    CV_TRY(hipHostMalloc((void **)&c->red_host, 4096 * sizeof(double), hipHostMallocDefault));
    c->V = nullptr; c->Z = nullptr;
  }
  CV_TRY(hipStreamSynchronize(c->stream));
  *out = c;
  return CV_OK;
fail:
#undef CV_TRY
  (void)hipStreamSynchronize(c->stream);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->red_host) (void)hipHostFree(c->red_host);
  for (auto e : c->step_ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return rc;
}

extern "C" int cv_ctx_destroy(cv_ctx *c) {
  if (!c) return CV_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->red_host) (void)hipHostFree(c->red_host);
  for (auto e : c->prof.ev) (void)hipEventDestroy(e);
  for (auto e : c->step_ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return CV_OK;
}

// ---------------------------------------------------------------------
// field exchange: host AoS (Cmpnts) <-> device planar
// ---------------------------------------------------------------------
static double *const *field_ptr(cv_ctx *c, int f, int level, int *dof) {
  static thread_local double *tmp[3];
  CvLevel &L = c->lev[level];
  CvMetrics &M = c->m;
  int fine = (level == c->nlev - 1);
  *dof = 1;
  switch (f) {
    case CV_F_UCONT: *dof = 3; return fine ? c->ucont : nullptr;
    case CV_F_UCONT_O: *dof = 3; return fine ? c->ucont_o : nullptr;
    case CV_F_UCAT: *dof = 3; return fine ? c->ucat : nullptr;
    case CV_F_RHS: *dof = 3; return fine ? c->rhs : nullptr;
    case CV_F_P: tmp[0] = c->p; return fine ? tmp : nullptr;
    case CV_F_PHI: tmp[0] = c->phi; return fine ? tmp : nullptr;
    case CV_F_B: tmp[0] = c->bvec; return fine ? tmp : nullptr;
    case CV_F_NVERT: tmp[0] = L.nvert; return tmp;
    case CV_F_AJ: tmp[0] = L.aj; return tmp;
    case CV_F_CSI: *dof = 3; return fine ? M.csi : nullptr;
    case CV_F_ETA: *dof = 3; return fine ? M.eta : nullptr;
    case CV_F_ZET: *dof = 3; return fine ? M.zet : nullptr;
    case CV_F_ICSI: *dof = 3; return fine ? M.icsi : nullptr;
    case CV_F_IETA: *dof = 3; return fine ? M.ieta : nullptr;
    case CV_F_IZET: *dof = 3; return fine ? M.izet : nullptr;
    case CV_F_JCSI: *dof = 3; return fine ? M.jcsi : nullptr;
    case CV_F_JETA: *dof = 3; return fine ? M.jeta : nullptr;
    case CV_F_JZET: *dof = 3; return fine ? M.jzet : nullptr;
    case CV_F_KCSI: *dof = 3; return fine ? M.kcsi : nullptr;
    case CV_F_KETA: *dof = 3; return fine ? M.keta : nullptr;
    case CV_F_KZET: *dof = 3; return fine ? M.kzet : nullptr;
    case CV_F_IAJ: tmp[0] = M.iaj; return fine ? tmp : nullptr;
    case CV_F_JAJ: tmp[0] = M.jaj; return fine ? tmp : nullptr;
    case CV_F_KAJ: tmp[0] = M.kaj; return fine ? tmp : nullptr;
    default: return nullptr;
  }
}

__global__ void k_aos_to_planar(const double *aos, double *x, double *y, double *z, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    x[t] = aos[3 * t]; y[t] = aos[3 * t + 1]; z[t] = aos[3 * t + 2];
  }
}
__global__ void k_planar_to_aos(double *aos, const double *x, const double *y, const double *z, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    aos[3 * t] = x[t]; aos[3 * t + 1] = y[t]; aos[3 * t + 2] = z[t];
  }
}

extern "C" int cv_get_field(cv_ctx *c, int f, int level, double *out) {
  if (!c || !out) return CV_ERR_ARG;
  if (level < 0) level = c->nlev - 1;
  if (level >= c->nlev) return CV_ERR_ARG;
  int dof;
  double *const *ptr = field_ptr(c, f, level, &dof);
  if (!ptr) { cv_set_error("cv_get_field: field %d not available on level %d", f, level); return CV_ERR_ARG; }
  long n = c->lev[level].d.n;
  if (dof == 1) {
    CV_CHECK(hipMemcpyAsync(out, ptr[0], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else {
    hipLaunchKernelGGL(k_planar_to_aos, grid1(n, 256), dim3(256), 0, c->stream, c->stage, ptr[0], ptr[1], ptr[2], n);
    CV_CHECK(hipMemcpyAsync(out, c->stage, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  CV_CHECK(hipStreamSynchronize(c->stream));
  return CV_OK;
}

extern "C" int cv_put_field(cv_ctx *c, int f, int level, const double *in) {
  if (!c || !in) return CV_ERR_ARG;
  if (f == CV_F_UCONT) c->steps_done = 0;     // loaded fluxes start a run (BDF1 first)
  if (level < 0) level = c->nlev - 1;
  if (level >= c->nlev) return CV_ERR_ARG;
  int dof;
  double *const *ptr = field_ptr(c, f, level, &dof);
  if (!ptr) { cv_set_error("cv_put_field: field %d not available on level %d", f, level); return CV_ERR_ARG; }
  long n = c->lev[level].d.n;
  if (dof == 1) {
    CV_CHECK(hipMemcpyAsync(ptr[0], in, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else {
    CV_CHECK(hipMemcpyAsync(c->stage, in, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_aos_to_planar, grid1(n, 256), dim3(256), 0, c->stream, (const double *)c->stage,
                       ptr[0], ptr[1], ptr[2], n);
  }
  CV_CHECK(hipStreamSynchronize(c->stream));
  return CV_OK;
}

// ---------------------------------------------------------------------
// kernel timing (device events on the context stream)
// ---------------------------------------------------------------------
#define CV_PROF_POOL 256

CvProfScope::CvProfScope(cv_ctx *c_, int kind, int level, int prec, double bytes) : c(c_), on(false) {
  CvProf &P = c->prof;
  if (P.kind == CV_PROF_NONE || P.kind != kind) return;
  const int lev = P.level < 0 ? c->nlev - 1 : P.level;
  if (lev != level || (P.prec >= 0 && P.prec != prec)) return;
  if (P.pending == CV_PROF_POOL && cv_prof_drain(c) != CV_OK) return;
  if (hipEventRecord(P.ev[2 * P.pending], c->stream) != hipSuccess) return;
  on = true;
  P.bytes += bytes;
}

CvProfScope::~CvProfScope() {
  if (!on) return;
  CvProf &P = c->prof;
  (void)hipEventRecord(P.ev[2 * P.pending + 1], c->stream);
  P.pending++;
  P.launches++;
}

int cv_prof_drain(cv_ctx *c) {
  CvProf &P = c->prof;
  if (P.pending == 0) return CV_OK;
  CV_CHECK(hipEventSynchronize(P.ev[2 * P.pending - 1]));
  for (int q = 0; q < P.pending; q++) {
    float ms = 0.f;
    CV_CHECK(hipEventElapsedTime(&ms, P.ev[2 * q], P.ev[2 * q + 1]));
    P.ms += ms;
  }
  P.pending = 0;
  return CV_OK;
}

extern "C" int cv_prof_select(cv_ctx *c, int kernel, int level, int precision) {
  if (!c || kernel < CV_PROF_NONE || kernel > CV_PROF_ORTH || level >= c->nlev) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(c->device));
  int rc = cv_prof_drain(c);
  if (rc) return rc;
  CvProf &P = c->prof;
  if (P.ev.empty()) {
    P.ev.resize(2 * CV_PROF_POOL);
    for (auto &e : P.ev) CV_CHECK(hipEventCreate(&e));
  }
  P.kind = kernel; P.level = level; P.prec = precision;
  P.launches = 0; P.ms = 0.0; P.bytes = 0.0;
  return CV_OK;
}

extern "C" int cv_prof_read(cv_ctx *c, long *launches, double *total_ms, double *bytes, int reset) {
  if (!c) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(c->device));
  int rc = cv_prof_drain(c);
  if (rc) return rc;
  CvProf &P = c->prof;
  if (launches) *launches = P.launches;
  if (total_ms) *total_ms = P.ms;
  if (bytes) *bytes = P.bytes;
  if (reset) { P.launches = 0; P.ms = 0.0; P.bytes = 0.0; }
  return CV_OK;
}
