// cv_batch_solve.h -- the batched pressure Poisson solve (PoissonSolver_MG,
// poisson.c:1775-2017, for many independent cases at once).  Included at the
// end of cv_poisson.hip: it runs that file's device bodies.
//
// One launch per operation covers one multigrid depth (0 = finest) of every
// case in the batch.  A block table maps each block of the launch to
// (case, block of that case's own launch shape), and the block runs the
// single-case body on that case's arrays.  FGMRES runs in lockstep over the
// cases with one host synchronisation per iteration; converged cases are
// masked out.  The preconditioner and operator apply of one iteration are
// captured once in a hipGraph and replayed: the Krylov index and the mask
// of active cases live in device memory, so the graph never changes.  Every
// case gets the launch shapes and reduction partitions it gets alone, so
// its result is bit-identical to cv_poisson_solve on that case.
#pragma once
#include <algorithm>
#include <cstring>

namespace {

#define BMAXC 64      // cases per batch (one bit each in a 64-bit mask)
#ifndef CV_BATCH_LEGS_DEFAULT
// first depth of the persistent V-cycle legs; off (-1): measured slower than
// one launch per operation on the 64-case ensemble at every starting depth
#define CV_BATCH_LEGS_DEFAULT -1
#endif
#define BSC 48        // per-case host-set scalars: [0] 1/beta, [1] 1/hn, [2] m, [8..] y

struct BLev {
  OpV<float> Af;
  OpV<double> Ad;
  float *xf, *bf, *rf;
  double *xd, *bd, *rd;
  const double *nv;
  const float *nwf;     // max(0, 1 - nvert) in FP32 (restriction weights)
  const uint32_t *fl;
  Idx I;
  int isc, jsc, ksc;
  long n;
  double inv_nf;
  int sgx, sgy, sgz;   // stencil tiles in i, j, k
  int kc;              // k planes per stencil tile
  int lgx, cs;         // line-solve blocks (64 lines each); rows per thread of the line solve
  int nvb, nrd;        // blocks of an elementwise pass and of a reduction
};

struct BCase {
  BLev lev[CV_MAXLEV];   // by depth, 0 = finest
  double *V, *Z, *w, *phi, *rhs;
  double *part, *dsc, *bsc;
  long n;
  int nlev;
  // the step around the solve: Ucont, the cell face-area vectors, p
  double *uc[3];
  const double *csi[3], *eta[3], *zet[3];
  double *p;
  double dt, st;
};

struct BCtl {
  unsigned long long act;   // cases in this FGMRES iteration
  int jj;                   // Krylov index
  int pad;
};

// vector slots of a case: level vectors (any precision) and, in FP64, the
// fine-level FGMRES vectors
enum { BX = 0, BB = 1, BR = 2, BZJ = 3, BW = 4, BPHI = 5, BRHS = 6, BVJ = 7, BVJ1 = 8 };

template <class T>
__device__ inline T *bvec(const BCase &C, int d, int s, int jj);
template <>
__device__ inline float *bvec<float>(const BCase &C, int d, int s, int jj) {
  // BZJ: the FP32 preconditioned vector Z_jj of the mixed solve, which is
  // also the x of the finest V-cycle level (the V-cycle writes M^-1 V_jj in
  // place; only the mixed solve runs an FP32 V-cycle)
  float *const zj = (float *)C.Z + (long)jj * C.n;
  return s == BX ? (d == 0 ? zj : C.lev[d].xf) : s == BB ? C.lev[d].bf : s == BZJ ? zj : C.lev[d].rf;
}
template <>
__device__ inline double *bvec<double>(const BCase &C, int d, int s, int jj) {
  switch (s) {
    case BX: return C.lev[d].xd;
    case BB: return C.lev[d].bd;
    case BR: return C.lev[d].rd;
    case BZJ: return C.Z + (long)jj * C.n;
    case BW: return C.w;
    case BPHI: return C.phi;
    case BRHS: return C.rhs;
    case BVJ: return C.V + (long)jj * C.n;
    default: return C.V + (long)(jj + 1) * C.n;
  }
}
template <class T>
__device__ inline OpV<T> bop(const BCase &C, int d);
template <>
__device__ inline OpV<float> bop<float>(const BCase &C, int d) { return C.lev[d].Af; }
template <>
__device__ inline OpV<double> bop<double>(const BCase &C, int d) { return C.lev[d].Ad; }

// block table entry: case << 24 | block within the case's launch shape;
// false when the case sits this launch out (the whole block leaves)
__device__ inline bool bblock(const uint32_t *tab, unsigned long long pm, const BCtl *ctl, int &c, int &lb) {
  const uint32_t e = tab[blockIdx.x];
  c = (int)(e >> 24);
  lb = (int)(e & 0xffffffu);
  return ((pm & ctl->act) >> c) & 1ull;
}

#define BK_ARGS const BCase *__restrict__ cs, const uint32_t *__restrict__ tab, unsigned long long pm, \
                const BCtl *__restrict__ ctl
#define BK_HEAD                                  \
  int c, lb;                                     \
  if (!bblock(tab, pm, ctl, c, lb)) return;      \
  const BCase &C = cs[c];

// TX: element type of the input vector (FP32 for the mixed solve's Z_jj
// under the FP64 apply)
template <class T, int MODE, bool HG, int TW, class TX = T>
__global__ void __launch_bounds__(TW * 4) kb_stencil(BK_ARGS, int d, int sx, int sb, int sy) {
  BK_HEAD
  const OpV<T> A = bop<T>(C, d);
  const int gx = C.lev[d].sgx, gy = C.lev[d].sgy, jj = ctl->jj;
  const int bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
  stencil_tile<T, MODE, 4, true, HG, TW, TX>(A, (const TX *)bvec<TX>(C, d, sx, jj),
                                     MODE >= 1 ? bvec<T>(C, d, sb, jj) : (const T *)nullptr, bvec<T>(C, d, sy, jj),
                                     A.dinv, (T)0, C.lev[d].kc, bx, by, bz);
}

// x += omega T^-1 r in place (zg: x = omega T^-1 r); PX: x loaded with r
// (launches of at most one block per CU)
template <class T, int CS, bool HL = false, bool PX = false>
__global__ void __launch_bounds__(1024) kb_yline(BK_ARGS, int d, int sr, T omega, int zg) {
  BK_HEAD
  const OpV<T> A = bop<T>(C, d);
  T *x = bvec<T>(C, d, BX, ctl->jj);
  yline_reg_tile<T, CS, HL, PX>(A.mx, A.my, A.mz, bvec<T>(C, d, sr, 0), A.ylo, A.yip, A.ycp, x, x,
                                HL ? (T)(omega * A.ysc) : omega, zg, lb, A.yh, A.yc);
}

// r of depth d -> b of depth d+1 (block table of depth d+1)
template <class T>
__global__ void kb_restrict(BK_ARGS, int d) {
  BK_HEAD
  const BLev &F = C.lev[d], &K = C.lev[d + 1];
  restrict_body<T>(K.I, F.I, K.isc, K.jsc, K.ksc, F.nv, F.nwf, K.fl, (const T *)bvec<T>(C, d, BR, 0),
                   bvec<T>(C, d + 1, BB, 0), K.n, (long)lb * BS + threadIdx.x, (long)K.nvb * BS);
}

// x of depth d += P x of depth d+1 (block table of depth d)
template <class T>
__global__ void kb_prolong(BK_ARGS, int d) {
  BK_HEAD
  const BLev &F = C.lev[d], &K = C.lev[d + 1];
  prolong_body<T>(F.I, K.I, K.isc, K.jsc, K.ksc, K.nv, F.fl, (const T *)bvec<T>(C, d + 1, BX, ctl->jj),
                  bvec<T>(C, d, BX, ctl->jj), 1, F.n, (long)lb * BS + threadIdx.x, (long)F.nvb * BS);
}

template <class T>
__global__ void kb_fluid_sum(BK_ARGS, int d, int sx) {
  BK_HEAD
  fluid_sum_body<T>((const T *)bvec<T>(C, d, sx, ctl->jj), C.lev[d].fl, C.lev[d].n, C.part, lb, C.lev[d].nrd);
}

template <class T>
__global__ void kb_fluid_shift(BK_ARGS, int d, int sx) {
  BK_HEAD
  fluid_shift_body<T>(bvec<T>(C, d, sx, ctl->jj), C.lev[d].fl, C.dsc + 100, C.lev[d].inv_nf, C.lev[d].n,
                      (long)lb * BS + threadIdx.x, (long)C.lev[d].nvb * BS);
}

// dsc[slot + q] = sum of the case's partials of a depth-d reduction; grid
// (cases, quantities)
__global__ void kb_final(const BCase *__restrict__ cs, unsigned long long pm, const BCtl *__restrict__ ctl, int d,
                         int slot) {
  const int c = blockIdx.x;
  if (!(((pm & ctl->act) >> c) & 1ull)) return;
  const BCase &C = cs[c];
  final_sum_body(C.part, C.lev[d].nrd, C.dsc + slot, blockIdx.y);
}

template <class T>
__global__ void kb_zero(BK_ARGS, int d, int sx) {
  BK_HEAD
  T *x = bvec<T>(C, d, sx, ctl->jj);
  const long n = C.lev[d].n;
  for (long t = (long)lb * BS + threadIdx.x; t < n; t += (long)C.lev[d].nvb * BS) x[t] = (T)0;
}

// dir 0: V_jj -> b of the finest V-cycle level; dir 1: x of that level -> Z_jj
// (the FP64 solve; the mixed solve's scale writes that b, its V-cycle Z_jj)
template <class T>
__global__ void kb_conv(BK_ARGS, int dir) {
  BK_HEAD
  const int jj = ctl->jj;
  const long n = C.n, st = (long)C.lev[0].nvb * BS;
  if (dir == 0) {
    const double *s = bvec<double>(C, 0, BVJ, jj);
    T *o = bvec<T>(C, 0, BB, jj);
    for (long t = (long)lb * BS + threadIdx.x; t < n; t += st) o[t] = (T)s[t];
  } else {   // the V-cycle's result to Z_jj (FP32 in the mixed solve)
    const T *s = bvec<T>(C, 0, BX, jj);
    T *o = bvec<T>(C, 0, BZJ, jj);
    for (long t = (long)lb * BS + threadIdx.x; t < n; t += st) o[t] = s[t];
  }
}

__global__ void kb_copy(BK_ARGS, int sdst, int ssrc) {
  BK_HEAD
  const int jj = ctl->jj;
  double *o = bvec<double>(C, 0, sdst, jj);
  const double *s = bvec<double>(C, 0, ssrc, jj);
  for (long t = (long)lb * BS + threadIdx.x; t < C.n; t += (long)C.lev[0].nvb * BS) o[t] = s[t];
}

// dst = src * bsc[bslot]; f32: also the FP32 copy the mixed V-cycle takes
// as its right-hand side (the b of the finest level)
__global__ void kb_scale(BK_ARGS, int sdst, int ssrc, int bslot, int f32) {
  BK_HEAD
  const int jj = ctl->jj;
  double *const dst = bvec<double>(C, 0, sdst, jj);
  const double *const src = bvec<double>(C, 0, ssrc, jj);
  const double a = C.bsc[bslot];
  const long t0 = (long)lb * BS + threadIdx.x, st = (long)C.lev[0].nvb * BS;
  if (!f32) {
    scale_to_body(dst, src, a, C.n, t0, st);
    return;
  }
  float *const bf = C.lev[0].bf;
  for (long t = t0; t < C.n; t += st) {
    const double v = src[t] * a;
    dst[t] = v;
    bf[t] = (float)v;
  }
}

template <class TZ>
__global__ void kb_xupdate(BK_ARGS) {
  BK_HEAD
  xupdate_body<TZ>(C.phi, (const TZ *)C.Z, C.n, C.bsc + 8, (int)C.bsc[2], C.n, (long)lb * BS + threadIdx.x,
                   (long)C.lev[0].nvb * BS);
}

__global__ void kb_sumsq(BK_ARGS, int sx) {
  BK_HEAD
  sumsq_body(bvec<double>(C, 0, sx, ctl->jj), C.n, C.part, lb, C.lev[0].nrd);
}

template <int NV>
__global__ void __launch_bounds__(BS) kb_mdot(BK_ARGS) {
  BK_HEAD
  mdot_body<NV>(C.w, C.V, C.n, C.n, C.part, lb, C.lev[0].nrd);
}

template <int NV>
__global__ void __launch_bounds__(BS) kb_orth(BK_ARGS) {
  BK_HEAD
  orth_body<NV>(C.w, C.V, C.n, C.dsc, C.n, C.part, lb, C.lev[0].nrd);
}

typedef void (*BRedFn)(const BCase *, const uint32_t *, unsigned long long, const BCtl *);
#define CV_BMDOT_ENTRY(N) kb_mdot<N>,
#define CV_BORTH_ENTRY(N) kb_orth<N>,
static const BRedFn bmdot_tbl[] = {CV_NV_LIST(CV_BMDOT_ENTRY)};
static const BRedFn borth_tbl[] = {CV_NV_LIST(CV_BORTH_ENTRY)};

// ---- the step around the solve (no FGMRES mask) ----
#define BK2_ARGS const BCase *__restrict__ cs, const uint32_t *__restrict__ tab, unsigned long long pm
#define BK2_HEAD                                                   \
  const uint32_t e_ = tab[blockIdx.x];                             \
  const int c = (int)(e_ >> 24), lb = (int)(e_ & 0xffffffu);       \
  if (!((pm >> c) & 1ull)) return;                                 \
  const BCase &C = cs[c];
#define BK2_GS (long)lb * BS + threadIdx.x, (long)C.lev[0].nvb * BS

// VolumeFlux_rev: mode 0 sums IB flux and area, mode 1 applies dsc[108]
__global__ void kb_ibm_flux(BK2_ARGS, int mode) {
  BK2_HEAD
  const BLev &L = C.lev[0];
  ibm_flux_body(L.I, L.nv, C.uc[0], C.uc[1], C.uc[2], C.csi[0], C.csi[1], C.csi[2], C.eta[0], C.eta[1], C.eta[2],
                C.zet[0], C.zet[1], C.zet[2], mode, C.dsc + 108, C.n, C.part, lb, L.nrd);
}
__global__ void kb_ibm_corr(const BCase *__restrict__ cs, unsigned long long pm) {
  const int c = blockIdx.x;
  if (!((pm >> c) & 1ull) || threadIdx.x) return;
  ibm_corr_body(cs[c].dsc + 112, cs[c].dsc + 108);
}
__global__ void kb_poisson_rhs(BK2_ARGS) {
  BK2_HEAD
  const BLev &L = C.lev[0];
  poisson_rhs_body(L.I, C.uc[0], C.uc[1], C.uc[2], L.Ad.aj, L.nv, C.rhs, C.dt, C.st, C.n, C.part, lb, L.nrd);
}
// dsc[slot + q] = sum of the fine-level partials, grid (cases, quantities)
__global__ void kb_final2(const BCase *__restrict__ cs, unsigned long long pm, int slot) {
  const int c = blockIdx.x;
  if (!((pm >> c) & 1ull)) return;
  const BCase &C = cs[c];
  final_sum_body(C.part, C.lev[0].nrd, C.dsc + slot, blockIdx.y);
}
__global__ void kb_p_add(BK2_ARGS) {
  BK2_HEAD
  p_add_body(C.lev[0].I, C.p, C.phi, C.n, BK2_GS);
}
// UpdatePressure's periodic planes: axis 0 (x) or 2 (z), interior or full range
__global__ void kb_pplane(BK2_ARGS, int axis, int full) {
  BK2_HEAD
  const Idx I = C.lev[0].I;
  if (axis == 0)
    pplane_x_body(I, C.p, C.phi, full ? 0 : 1, full ? I.my : I.my - 1, full ? 0 : 1, full ? I.mz : I.mz - 1, BK2_GS);
  else
    pplane_z_body(I, C.p, C.phi, full ? 0 : 1, full ? I.mx : I.mx - 1, full ? 0 : 1, full ? I.my : I.my - 1, BK2_GS);
}
__global__ void kb_projection(BK2_ARGS) {
  BK2_HEAD
  const BLev &L = C.lev[0];
  const FaceMet Mi{{L.Ad.g[0], L.Ad.g[1], L.Ad.g[2]}};
  const FaceMet Mj{{L.Ad.g[3], L.Ad.g[4], L.Ad.g[5]}};
  const FaceMet Mk{{L.Ad.g[6], L.Ad.g[7], L.Ad.g[8]}};
  projection_body(L.I, L.Ad.xper, L.Ad.zper, C.phi, L.nv, L.fl, Mi, Mj, Mk, C.uc[0], C.uc[1], C.uc[2], C.dt, C.st,
                  C.n, BK2_GS);
}

// dsc[0..BSC) of every case -> out[c*BSC ..]
__global__ void kb_gather(const BCase *__restrict__ cs, double *out) {
  const int c = blockIdx.x;
  if (threadIdx.x < BSC) out[(long)c * BSC + threadIdx.x] = cs[c].dsc[threadIdx.x];
}

}  // namespace

// ====================================================================
// host side
// ====================================================================
struct BTab {
  long off = 0;
  int len = 0;
};

// per-case state of the lockstep FGMRES (the single-case loop's locals)
struct BFg {
  std::vector<double> H, cs, sn, g, y;
  double bnorm = 0, tol = 0, beta = 0, rnorm = 0, hn2 = 0, wn2 = 0;
  int its = 0, jn = 0;
};

// Divergence (main.c:469-593) of every case at once: block (b, c) reduces
// case c's share b of nb into the case's partials, then one block per case
// reduces those with the same order rule (largest value, then smallest
// index) into out[2c], out[2c+1]; the result equals cv_divergence_dev's
struct BDiv {
  Idx I;
  const double *ux, *uy, *uz, *aj, *nv;
  long n;
  double *part;
};
__global__ void kbd_divmax(const BDiv *cs, int nb) {
  const BDiv &C = cs[blockIdx.y];
  divmax_block(C.I, C.ux, C.uy, C.uz, C.aj, C.nv, C.n, C.part, C.part + nb, blockIdx.x, nb);
}
__global__ void kbd_final(const BDiv *cs, int nb, double *out) {
  __shared__ double sv[BS];
  __shared__ long si[BS];
  const BDiv &C = cs[blockIdx.x];
  double best = -1.0;
  long bi = 0;
  for (int b = threadIdx.x; b < nb; b += BS) {
    const double v = C.part[b];
    const long ii = (long)C.part[nb + b];
    if (div_beats(v, ii, best, bi)) { best = v; bi = ii; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double v2 = sv[threadIdx.x + s];
      const long i2 = si[threadIdx.x + s];
      if (div_beats(v2, i2, sv[threadIdx.x], si[threadIdx.x])) { sv[threadIdx.x] = v2; si[threadIdx.x] = i2; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[2 * blockIdx.x] = sv[0]; out[2 * blockIdx.x + 1] = (double)si[0]; }
}

struct BLegs;

struct cv_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0;
  std::vector<cv_ctx *> cs;
  std::vector<int> ib_gen;            // each case's IB generation at cv_batch_create
  cv_solver_opts opt{};
  int maxlev = 0;
  BCase *dcases = nullptr;
  BCtl *dctl = nullptr, *hctl = nullptr;
  double *dbsc = nullptr, *hbsc = nullptr, *dgat = nullptr, *hgat = nullptr;
  BDiv *ddiv = nullptr;                       // per-case Divergence operands
  uint32_t *dtab = nullptr;
  BTab st[CV_MAXLEV], vc[CV_MAXLEV], rd[CV_MAXLEV], ln[CV_MAXLEV][3];
  // the same tables restricted to the cases that go deeper [0] or bottom out [1] at the depth
  BTab st_v[CV_MAXLEV][2], vc_v[CV_MAXLEV][2], rd_v[CV_MAXLEV][2], ln_v[CV_MAXLEV][2][3];
  int lnc[CV_MAXLEV][3];
  unsigned long long all = 0, has[CV_MAXLEV], inner[CV_MAXLEV], coarse[CV_MAXLEV], lgm[CV_MAXLEV][3];
  hipGraphExec_t gexec = nullptr;
  int use_graph = 1;
  std::vector<hipGraphExec_t> gpre, gpost;   // per-case graphs of the work around the solve
  CvMomBatch *mom = nullptr;                  // the momentum of all cases on the batch stream
  int use_mom = 1;
  int tw = 32;                                // lanes per stencil tile row (32 or 64)
  unsigned long long xper = 0, zper = 0;
  std::vector<BFg> fg;
  std::vector<int> rc;
  std::vector<hipEvent_t> ev_case;
  hipEvent_t ev_solved = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
  BLegs *legs = nullptr;                      // the V-cycle from a depth down as persistent legs
  int *herr = nullptr;                        // their error flag, read back with each gather
};

static void b_upload_ctl(cv_batch *B, unsigned long long act, int jj) {
  B->hctl->act = act;
  B->hctl->jj = jj;
  (void)hipMemcpyAsync(B->dctl, B->hctl, sizeof(BCtl), hipMemcpyHostToDevice, B->stream);
}
static void b_upload_bsc(cv_batch *B) {
  (void)hipMemcpyAsync(B->dbsc, B->hbsc, (size_t)B->n * BSC * sizeof(double), hipMemcpyHostToDevice, B->stream);
}
// dsc[0..BSC) of every case to the host, then wait
static int b_legs_error(cv_batch *B);
static int b_gather(cv_batch *B) {
  hipLaunchKernelGGL(kb_gather, dim3((unsigned)B->n), dim3(64), 0, B->stream, (const BCase *)B->dcases, B->dgat);
  CV_CHECK(hipMemcpyAsync(B->hgat, B->dgat, (size_t)B->n * BSC * sizeof(double), hipMemcpyDeviceToHost, B->stream));
  int rc = b_legs_error(B);
  if (rc) return rc;
  CV_CHECK(hipStreamSynchronize(B->stream));
  if (B->legs && *B->herr) {
    cv_set_error("cv_batch: a persistent V-cycle leg stopped at its wait limit");
    return CV_ERR_HIP;
  }
  return CV_OK;
}

#define BTAB(t) B->dtab + (t).off

// which block table a launch at depth d with case mask pm needs: -1 every
// case of the depth, 0 only cases that go deeper, 1 only cases whose
// coarsest level is d (fewer idle blocks in the coarse sweeps)
static int b_variant(const cv_batch *B, int d, unsigned long long pm) {
  const unsigned long long m = pm & B->has[d];
  if (m == B->has[d]) return -1;
  if ((m & ~B->inner[d]) == 0) return 0;
  if ((m & ~B->coarse[d]) == 0) return 1;
  return -1;
}
#define BPICK(full, var, d, pm) (b_variant(B, d, pm) < 0 ? (full)[d] : (var)[d][b_variant(B, d, pm)])
#define BLAUNCH_OK(t, pm) ((t).len > 0 && (pm) != 0)

template <class T, int MODE, class TX = T>
static void b_stencil(cv_batch *B, int d, unsigned long long pm, int sx, int sb, int sy) {
  const BTab &t = BPICK(B->st, B->st_v, d, pm);
  if (!BLAUNCH_OK(t, pm)) return;
  const bool hg = sizeof(T) == 4 && B->opt.precision == CV_MIXED16;
  const BCase *cs = B->dcases;
  const uint32_t *tb = BTAB(t);
  const BCtl *ctl = B->dctl;
  if (B->tw == 32) {
    if (hg)
      hipLaunchKernelGGL((kb_stencil<T, MODE, sizeof(T) == 4, 32, TX>), dim3((unsigned)t.len), dim3(32, 4), 0, B->stream, cs,
                         tb, pm, ctl, d, sx, sb, sy);
    else
      hipLaunchKernelGGL((kb_stencil<T, MODE, false, 32, TX>), dim3((unsigned)t.len), dim3(32, 4), 0, B->stream, cs, tb, pm,
                         ctl, d, sx, sb, sy);
  } else {
    if (hg)
      hipLaunchKernelGGL((kb_stencil<T, MODE, sizeof(T) == 4, 64, TX>), dim3((unsigned)t.len), dim3(64, 4), 0, B->stream, cs,
                         tb, pm, ctl, d, sx, sb, sy);
    else
      hipLaunchKernelGGL((kb_stencil<T, MODE, false, 64, TX>), dim3((unsigned)t.len), dim3(64, 4), 0, B->stream, cs, tb, pm,
                         ctl, d, sx, sb, sy);
  }
}

template <class T>
static void b_line(cv_batch *B, int d, unsigned long long pm, int sr, T om, bool zg) {
  const int var = b_variant(B, d, pm);
  for (int g = 0; g < 3; g++) {
    const BTab &t = var < 0 ? B->ln[d][g] : B->ln_v[d][var][g];
    const unsigned long long m = pm & B->lgm[d][g];
    if (!BLAUNCH_OK(t, m)) continue;
    const dim3 blk(64, (unsigned)B->lnc[d][g]);
    const BCase *cs = B->dcases;
    const uint32_t *tb = BTAB(t);
    const BCtl *ctl = B->dctl;
    if constexpr (sizeof(T) == 4) {
      if (B->opt.precision == CV_MIXED16 && g < 2) {      // FP16 line factors (8 or 16 rows per thread)
        const bool px = t.len <= CV_PX_BLOCKS;
        if (g == 0 && px)
          hipLaunchKernelGGL((kb_yline<T, 8, true, true>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m, ctl, d,
                             sr, om, zg ? 1 : 0);
        else if (g == 0)
          hipLaunchKernelGGL((kb_yline<T, 8, true>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m, ctl, d, sr, om,
                             zg ? 1 : 0);
        else if (px)
          hipLaunchKernelGGL((kb_yline<T, 16, true, true>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m, ctl, d,
                             sr, om, zg ? 1 : 0);
        else
          hipLaunchKernelGGL((kb_yline<T, 16, true>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m, ctl, d, sr, om,
                             zg ? 1 : 0);
        continue;
      }
    }
    if (g == 0)
      hipLaunchKernelGGL((kb_yline<T, 8>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m, ctl, d, sr, om, zg ? 1 : 0);
    else if (g == 1)
      hipLaunchKernelGGL((kb_yline<T, 16>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m, ctl, d, sr, om, zg ? 1 : 0);
    else
      hipLaunchKernelGGL((kb_yline<T, sizeof(T) == 8 ? 16 : 32>), dim3((unsigned)t.len), blk, 0, B->stream, cs, tb, m,
                         ctl, d, sr, om, zg ? 1 : 0);
  }
}

template <class T>
static void b_zero(cv_batch *B, int d, unsigned long long pm, int sx) {
  const BTab &t = BPICK(B->vc, B->vc_v, d, pm);
  if (!BLAUNCH_OK(t, pm)) return;
  hipLaunchKernelGGL(kb_zero<T>, dim3((unsigned)t.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(t), pm, (const BCtl *)B->dctl, d, sx);
}

static void b_final(cv_batch *B, int d, unsigned long long pm, int slot, int nq) {
  hipLaunchKernelGGL(kb_final, dim3((unsigned)B->n, (unsigned)nq), dim3(BS), 0, B->stream, (const BCase *)B->dcases, pm,
                     (const BCtl *)B->dctl, d, slot);
}

template <class T>
static void b_nullspace(cv_batch *B, int d, unsigned long long pm, int sx) {
  const BTab &tr = BPICK(B->rd, B->rd_v, d, pm), &tv = BPICK(B->vc, B->vc_v, d, pm);
  if (!BLAUNCH_OK(tr, pm)) return;
  hipLaunchKernelGGL(kb_fluid_sum<T>, dim3((unsigned)tr.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(tr), pm, (const BCtl *)B->dctl, d, sx);
  b_final(B, d, pm, 100, 1);
  hipLaunchKernelGGL(kb_fluid_shift<T>, dim3((unsigned)tv.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(tv), pm, (const BCtl *)B->dctl, d, sx);
}

#include "cv_batch_legs.h"

// the V-cycle of depth d for every case that has it; the cases whose
// coarsest level is at depth d run the coarse sweeps and the nullspace rule
template <class T>
static void b_vcycle(cv_batch *B, int d) {
  if (B->legs && d == B->legs->d0) {
    b_vcycle_legs<T>(B);
    return;
  }
  const cv_solver_opts &o = B->opt;
  const T om = (T)o.omega;
  const unsigned long long in = B->inner[d], co = B->coarse[d];
  const int npre = o.pre_sweeps, npost = o.post_sweeps, nco = o.coarse_sweeps;
  const unsigned long long z = (npre == 0 ? in : 0) | (nco == 0 ? co : 0);
  if (z) b_zero<T>(B, d, z, BX);
  const int nsw = std::max(in ? npre : 0, co ? nco : 0);
  for (int s = 0; s < nsw; s++) {
    const unsigned long long m = (s < npre ? in : 0) | (s < nco ? co : 0);
    if (s > 0) b_stencil<T, 1>(B, d, m, BX, BB, BR);
    b_line<T>(B, d, m, s == 0 ? BB : BR, om, s == 0);
  }
  if (co) b_nullspace<T>(B, d, co, BX);
  if (!in) return;
  b_stencil<T, 1>(B, d, in, BX, BB, BR);
  hipLaunchKernelGGL(kb_restrict<T>, dim3((unsigned)B->vc[d + 1].len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(B->vc[d + 1]), in, (const BCtl *)B->dctl, d);
  b_vcycle<T>(B, d + 1);
  {
    const BTab &tp = BPICK(B->vc, B->vc_v, d, in);
    hipLaunchKernelGGL(kb_prolong<T>, dim3((unsigned)tp.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                       (const uint32_t *)BTAB(tp), in, (const BCtl *)B->dctl, d);
  }
  for (int s = 0; s < npost; s++) {
    b_stencil<T, 1>(B, d, in, BX, BB, BR);
    b_line<T>(B, d, in, BR, om, false);
  }
}

// Z_jj = M^-1 V_jj (nullspace rule applied) and w = A Z_jj for the cases
// of the device mask
static void b_precond_apply(cv_batch *B) {
  const unsigned long long all = B->all;
  const BTab &v0 = B->vc[0];
  if (B->opt.precision != CV_FP64) {
    // Z_jj in FP32, as in the single-case solve: the scale that wrote V_jj
    // wrote its FP32 copy (the V-cycle's b), and the V-cycle's finest x is Z_jj
    (void)v0;
    b_vcycle<float>(B, 0);
    b_nullspace<float>(B, 0, all, BZJ);
    b_stencil<double, 0, float>(B, 0, all, BZJ, BZJ, BW);
  } else {
    hipLaunchKernelGGL(kb_conv<double>, dim3((unsigned)v0.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                       (const uint32_t *)BTAB(v0), all, (const BCtl *)B->dctl, 0);
    b_vcycle<double>(B, 0);
    hipLaunchKernelGGL(kb_conv<double>, dim3((unsigned)v0.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                       (const uint32_t *)BTAB(v0), all, (const BCtl *)B->dctl, 1);
    b_nullspace<double>(B, 0, all, BZJ);
    b_stencil<double, 0>(B, 0, all, BZJ, BZJ, BW);
  }
}

static void b_vec(cv_batch *B, void (*k)(BK_ARGS, int, int), unsigned long long pm, int a0, int a1) {
  const BTab &t = B->vc[0];
  if (!BLAUNCH_OK(t, pm)) return;
  hipLaunchKernelGGL(k, dim3((unsigned)t.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(t), pm, (const BCtl *)B->dctl, a0, a1);
}
static void b_copy(cv_batch *B, unsigned long long pm, int dst, int src) { b_vec(B, kb_copy, pm, dst, src); }
static void b_sumsq(cv_batch *B, unsigned long long pm, int sx) {
  const BTab &t = B->rd[0];
  if (!BLAUNCH_OK(t, pm)) return;
  hipLaunchKernelGGL(kb_sumsq, dim3((unsigned)t.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(t), pm, (const BCtl *)B->dctl, sx);
  b_final(B, 0, pm, 40, 1);
}
static void b_scale(cv_batch *B, unsigned long long pm, int dst, int src, int slot) {
  const BTab &t = B->vc[0];
  if (!BLAUNCH_OK(t, pm)) return;
  hipLaunchKernelGGL(kb_scale, dim3((unsigned)t.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(t), pm, (const BCtl *)B->dctl, dst, src, slot,
                     B->opt.precision != CV_FP64 ? 1 : 0);
}
// one classical Gram-Schmidt pass: h and ||w||^2 to dsc[0..nv], ||w - V h||^2 to dsc[36]
static void b_gs(cv_batch *B, unsigned long long pm, int nv) {
  const BTab &t = B->rd[0];
  if (!BLAUNCH_OK(t, pm)) return;
  hipLaunchKernelGGL(bmdot_tbl[nv - 1], dim3((unsigned)t.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(t), pm, (const BCtl *)B->dctl);
  b_final(B, 0, pm, 0, nv + 1);
  hipLaunchKernelGGL(borth_tbl[nv - 1], dim3((unsigned)t.len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                     (const uint32_t *)BTAB(t), pm, (const BCtl *)B->dctl);
  b_final(B, 0, pm, 36, 1);
}

static int b_run_precond(cv_batch *B) {
  if (B->use_graph) {
    if (!B->gexec) {
      hipGraph_t g = nullptr;
      CV_CHECK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeRelaxed));
      b_precond_apply(B);
      CV_CHECK(hipStreamEndCapture(B->stream, &g));
      const hipError_t e = hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        cv_set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        B->gexec = nullptr;
        return CV_ERR_HIP;
      }
    }
    CV_CHECK(hipGraphLaunch(B->gexec, B->stream));
  } else {
    b_precond_apply(B);
  }
  return CV_OK;
}

static inline unsigned long long bit(int c) { return 1ull << c; }

// lockstep FGMRES(m) over the cases: phi = A^-1 rhs per case (phi starts at
// 0), the single-case loop of fgmres() run for all cases at once
static int b_fgmres(cv_batch *B) {
  const int nc = B->n, m = B->opt.restart;
  const double *G = B->hgat;
  int rc;
  B->fg.assign(nc, BFg());
  for (auto &f : B->fg) { f.H.assign((m + 1) * m, 0.0); f.cs.assign(m, 0.0); f.sn.assign(m, 0.0); f.g.assign(m + 1, 0.0); f.y.assign(m, 0.0); }
  B->rc.assign(nc, CV_OK);
  // ||b||
  b_upload_ctl(B, B->all, 0);
  b_sumsq(B, B->all, BRHS);
  if ((rc = b_gather(B))) return rc;
  unsigned long long cyc = 0;
  for (int c = 0; c < nc; c++) {
    BFg &f = B->fg[c];
    f.bnorm = std::sqrt(G[c * BSC + 40]);
    f.tol = std::fmax(B->opt.rtol * f.bnorm, B->opt.atol);
    f.beta = f.rnorm = f.bnorm;
    // a case whose right-hand side is not finite leaves the solve; the
    // others go on (their arithmetic does not depend on it)
    if (!std::isfinite(f.bnorm)) { B->rc[c] = CV_ERR_NONFINITE; continue; }
    if (f.beta > f.tol && f.its < B->opt.max_it) cyc |= bit(c);
  }
  b_zero<double>(B, 0, B->all, BPHI);
  b_copy(B, B->all, BW, BRHS);
  while (cyc) {
    for (int c = 0; c < nc; c++)
      if (cyc & bit(c)) B->hbsc[c * BSC + 0] = 1.0 / B->fg[c].beta;
    b_upload_bsc(B);
    b_scale(B, cyc, BVJ, BW, 0);   // the device mask still holds every case of cyc
    for (int c = 0; c < nc; c++)
      if (cyc & bit(c)) {
        BFg &f = B->fg[c];
        std::fill(f.g.begin(), f.g.end(), 0.0);
        f.g[0] = f.beta;
        f.jn = 0;
      }
    unsigned long long act = cyc;
    for (int jj = 0; jj < m && act; jj++) {
      const int nv = jj + 1;
      b_upload_ctl(B, act, jj);
      if ((rc = b_run_precond(B))) return rc;
      b_gs(B, act, nv);
      if ((rc = b_gather(B))) return rc;
      unsigned long long second = 0;
      for (int c = 0; c < nc; c++) {
        if (!(act & bit(c))) continue;
        BFg &f = B->fg[c];
        f.wn2 = G[c * BSC + nv];
        f.hn2 = G[c * BSC + 36];
        for (int q = 0; q < nv; q++) f.H[q * m + jj] = G[c * BSC + q];
        if (f.hn2 < 1e-2 * f.wn2) second |= bit(c);   // DGKS second pass
      }
      if (second) {
        b_gs(B, second, nv);
        if ((rc = b_gather(B))) return rc;
        for (int c = 0; c < nc; c++) {
          if (!(second & bit(c))) continue;
          BFg &f = B->fg[c];
          for (int q = 0; q < nv; q++) f.H[q * m + jj] += G[c * BSC + q];
          f.hn2 = G[c * BSC + 36];
        }
      }
      unsigned long long scal = 0, keep = act;
      for (int c = 0; c < nc; c++) {
        if (!(act & bit(c))) continue;
        BFg &f = B->fg[c];
        const double hn = std::sqrt(std::fmax(f.hn2, 0.0));
        f.H[(jj + 1) * m + jj] = hn;
        if (!std::isfinite(hn)) {   // this case leaves the solve
          B->rc[c] = CV_ERR_NONFINITE;
          keep &= ~bit(c);
          cyc &= ~bit(c);
          continue;
        }
        if (hn > 0) { B->hbsc[c * BSC + 1] = 1.0 / hn; scal |= bit(c); }
        std::vector<double> &H = f.H;
        for (int q = 0; q < jj; q++) {
          const double a = H[q * m + jj], bb = H[(q + 1) * m + jj];
          H[q * m + jj] = f.cs[q] * a + f.sn[q] * bb;
          H[(q + 1) * m + jj] = -f.sn[q] * a + f.cs[q] * bb;
        }
        const double a = H[jj * m + jj], bb = H[(jj + 1) * m + jj];
        const double den = std::sqrt(a * a + bb * bb);
        f.cs[jj] = den == 0 ? 1.0 : a / den;
        f.sn[jj] = den == 0 ? 0.0 : bb / den;
        H[jj * m + jj] = f.cs[jj] * a + f.sn[jj] * bb;
        H[(jj + 1) * m + jj] = 0.0;
        f.g[jj + 1] = -f.sn[jj] * f.g[jj];
        f.g[jj] = f.cs[jj] * f.g[jj];
        f.its++;
        f.rnorm = std::fabs(f.g[jj + 1]);
        f.jn = jj + 1;
        if (f.rnorm <= f.tol || hn == 0 || f.its >= B->opt.max_it) keep &= ~bit(c);
      }
      if (scal) {
        b_upload_bsc(B);
        b_scale(B, scal, BVJ1, BW, 1);   // V_{jj+1} = w / hn; the device mask holds act
      }
      act = keep;
    }
    // x += Z y, then the true residual
    for (int c = 0; c < nc; c++) {
      if (!(cyc & bit(c))) continue;
      BFg &f = B->fg[c];
      const int jn = f.jn;
      for (int q = jn - 1; q >= 0; q--) {
        double s = f.g[q];
        for (int t = q + 1; t < jn; t++) s -= f.H[q * m + t] * f.y[t];
        f.y[q] = s / f.H[q * m + q];
      }
      for (int q = 0; q < jn; q++) B->hbsc[c * BSC + 8 + q] = f.y[q];
      B->hbsc[c * BSC + 2] = (double)jn;
    }
    b_upload_bsc(B);
    b_upload_ctl(B, cyc, 0);
    {
      const BTab &t = B->vc[0];
      hipLaunchKernelGGL(B->opt.precision != CV_FP64 ? kb_xupdate<float> : kb_xupdate<double>, dim3((unsigned)t.len),
                         dim3(BS), 0, B->stream, (const BCase *)B->dcases,
                         (const uint32_t *)BTAB(t), cyc, (const BCtl *)B->dctl);
    }
    b_stencil<double, 1>(B, 0, cyc, BPHI, BRHS, BW);
    b_sumsq(B, cyc, BW);
    if ((rc = b_gather(B))) return rc;
    unsigned long long next = 0;
    for (int c = 0; c < nc; c++) {
      if (!(cyc & bit(c))) continue;
      BFg &f = B->fg[c];
      f.beta = std::sqrt(G[c * BSC + 40]);
      f.rnorm = f.beta;
      if (!std::isfinite(f.beta)) { B->rc[c] = CV_ERR_NONFINITE; continue; }
      if (f.beta > f.tol && f.its < B->opt.max_it) next |= bit(c);
    }
    cyc = next;
  }
  for (int c = 0; c < nc; c++)
    if (B->rc[c] != CV_ERR_NONFINITE) B->rc[c] = B->fg[c].rnorm <= B->fg[c].tol ? CV_OK : CV_ERR_NOT_CONVERGED;
  // the nullspace rule on phi (as after the single-case solve)
  b_upload_ctl(B, B->all, 0);
  b_nullspace<double>(B, 0, B->all, BPHI);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// ---------------- batch set-up --------------------------------------
static void b_release(cv_batch *B) {
  if (!B) return;
  (void)hipSetDevice(B->device);
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
  if (B->mom) cv_mom_batch_destroy(B->mom);
  for (auto g : B->gpre)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto g : B->gpost)
    if (g) (void)hipGraphExecDestroy(g);
  legs_free(B->legs);
  if (B->herr) (void)hipHostFree(B->herr);
  if (B->dcases) (void)hipFree(B->dcases);
  if (B->dctl) (void)hipFree(B->dctl);
  if (B->dbsc) (void)hipFree(B->dbsc);
  if (B->ddiv) (void)hipFree(B->ddiv);
  if (B->dgat) (void)hipFree(B->dgat);
  if (B->dtab) (void)hipFree(B->dtab);
  if (B->hctl) (void)hipHostFree(B->hctl);
  if (B->hbsc) (void)hipHostFree(B->hbsc);
  if (B->hgat) (void)hipHostFree(B->hgat);
  for (auto e : B->ev_case) (void)hipEventDestroy(e);
  if (B->ev_solved) (void)hipEventDestroy(B->ev_solved);
  if (B->ev_t0) (void)hipEventDestroy(B->ev_t0);
  if (B->ev_t1) (void)hipEventDestroy(B->ev_t1);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
}

static bool same_opts(const cv_solver_opts &a, const cv_solver_opts &b) {
  return a.precision == b.precision && a.smoother == b.smoother && a.pre_sweeps == b.pre_sweeps &&
         a.post_sweeps == b.post_sweeps && a.coarse_sweeps == b.coarse_sweeps && a.omega == b.omega &&
         a.restart == b.restart && a.max_it == b.max_it && a.rtol == b.rtol && a.atol == b.atol;
}

static int line_cs(int rows) { return rows <= 128 ? 8 : rows <= 256 ? 16 : 32; }

extern "C" int cv_batch_create(cv_ctx *const *cases, int ncases, cv_batch **out) {
  if (!cases || !out || ncases < 1 || ncases > BMAXC) {
    cv_set_error("cv_batch_create: 1..%d cases", BMAXC);
    return CV_ERR_ARG;
  }
  *out = nullptr;
  const int dev = cases[0]->device;
  for (int c = 0; c < ncases; c++) {
    cv_ctx *x = cases[c];
    if (!x || x->device != dev) { cv_set_error("cv_batch_create: every case must live on one device"); return CV_ERR_ARG; }
    if (!x->geometry_set || !x->nvert_set) { cv_set_error("cv_batch_create: case %d lacks geometry/nvert", c); return CV_ERR_STATE; }
    if (!same_opts(x->opt, cases[0]->opt)) { cv_set_error("cv_batch_create: case %d has other solver options", c); return CV_ERR_ARG; }
    for (int q = 0; q < c; q++)
      if (cases[q] == x) { cv_set_error("cv_batch_create: case %d given twice", c); return CV_ERR_ARG; }
  }
  const cv_solver_opts &o = cases[0]->opt;
  if (o.smoother != CV_SMOOTH_YLINE) { cv_set_error("cv_batch_create: the batched solve uses the y-line smoother"); return CV_ERR_ARG; }
  CV_CHECK(hipSetDevice(dev));
  cv_batch *B = new cv_batch();
  B->device = dev;
  B->n = ncases;
  B->cs.assign(cases, cases + ncases);
  for (int c = 0; c < ncases; c++) B->ib_gen.push_back(cases[c]->ib_gen);
  B->opt = o;
  const char *ge = getenv("CV_BATCH_GRAPH");
  B->use_graph = ge ? atoi(ge) != 0 : 1;
  const char *gm = getenv("CV_BATCH_MOM");
  B->use_mom = gm ? atoi(gm) != 0 : 1;
  const char *gt = getenv("CV_BATCH_TW");       // stencil tile row width: 32 (measured best) or 64 lanes
  B->tw = gt && atoi(gt) == 64 ? 64 : 32;
  int rc = CV_OK;
  std::vector<BCase> hc(ncases);
  std::vector<uint32_t> tabs;
  auto add_tab = [&](BTab &t, const std::vector<std::pair<int, int>> &per) {
    t.off = (long)tabs.size();
    long len = 0;
    for (auto &p : per)
      for (int b = 0; b < p.second; b++) { tabs.push_back(((uint32_t)p.first << 24) | (uint32_t)b); len++; }
    t.len = (int)len;
  };
  for (int l = 0; l < CV_MAXLEV; l++) {
    B->has[l] = B->inner[l] = B->coarse[l] = 0;
    for (int g = 0; g < 3; g++) { B->lgm[l][g] = 0; B->lnc[l][g] = 1; }
  }
  for (int c = 0; c < ncases; c++) {
    cv_ctx *x = cases[c];
    CV_CHECK(hipSetDevice(dev));
    if ((rc = ensure_krylov(x))) { b_release(B); return rc; }
    BCase &bc = hc[c];
    memset(&bc, 0, sizeof(bc));
    bc.nlev = x->nlev;
    B->maxlev = std::max(B->maxlev, x->nlev);
    for (int d = 0; d < x->nlev; d++) {
      const int l = x->nlev - 1 - d;
      CvLevel &L = x->lev[l];
      BLev &b = bc.lev[d];
      if (o.precision == CV_MIXED16 && !L.gh) {
        cv_set_error("cv_batch_create: case %d has no FP16 coefficients", c);
        b_release(B);
        return CV_ERR_STATE;
      }
      b.Af = opview<float>(x, l);
      b.Ad = opview<double>(x, l);
      b.xf = L.xf; b.bf = L.bf; b.rf = L.rf;
      b.xd = L.x; b.bd = L.b; b.rd = L.r;
      b.nv = L.nvert; b.nwf = L.nwf; b.fl = L.flags;
      b.I = idx_of(L);
      b.isc = L.d.isc; b.jsc = L.d.jsc; b.ksc = L.d.ksc;
      b.n = L.d.n;
      b.inv_nf = x->nfluid[l] > 0 ? 1.0 / x->nfluid[l] : 0.0;
      b.kc = stencil_kchunk_batch(L.d.mz);
      b.sgx = (L.d.mx + B->tw - 2) / (B->tw - 1); b.sgy = (L.d.my + 3) / 4; b.sgz = (L.d.mz + b.kc - 1) / b.kc;
      b.lgx = (int)(((long)L.d.mx * L.d.mz + 63) / 64);
      const int rows = L.d.my - 2;
      const int rmax = o.precision != CV_FP64 ? 512 : 256;
      if (rows > rmax || rows < 1) {
        cv_set_error("cv_batch_create: case %d has %d rows per y-line (the batched line solve takes 1..%d)", c, rows, rmax);
        b_release(B);
        return CV_ERR_ARG;
      }
      b.cs = line_cs(rows);
      const long nv = (b.n + BS - 1) / BS;
      b.nvb = (int)std::min(nv, 1024L);
      b.nrd = nred(b.n);
      B->has[d] |= bit(c);
      if (d == x->nlev - 1) B->coarse[d] |= bit(c);
      else B->inner[d] |= bit(c);
      const int g = b.cs == 8 ? 0 : b.cs == 16 ? 1 : 2;
      B->lgm[d][g] |= bit(c);
      B->lnc[d][g] = std::max(B->lnc[d][g], (rows + b.cs - 1) / b.cs);
    }
    CvLevel &F = x->lev[x->nlev - 1];
    if (F.d.xper) B->xper |= bit(c);
    if (F.d.zper) B->zper |= bit(c);
    bc.V = x->V; bc.Z = x->Z; bc.w = x->w; bc.phi = x->phi; bc.rhs = x->bvec;
    bc.part = x->part; bc.dsc = x->dsc;
    bc.n = F.d.n;
    for (int q = 0; q < 3; q++) {
      bc.uc[q] = x->ucont[q];
      bc.csi[q] = x->m.csi[q]; bc.eta[q] = x->m.eta[q]; bc.zet[q] = x->m.zet[q];
    }
    bc.p = x->p;
    bc.dt = x->dt;
    bc.st = x->st;
    B->all |= bit(c);
  }
  for (int d = 0; d < B->maxlev; d++) {
    // v = -1: every case of the depth; 0: the cases that go deeper; 1: those that bottom out here
    for (int v = -1; v < 2; v++) {
      std::vector<std::pair<int, int>> st, vc, rd, ln[3];
      const unsigned long long m = v < 0 ? B->has[d] : v == 0 ? B->inner[d] : B->coarse[d];
      for (int c = 0; c < ncases; c++) {
        if (!(m & bit(c))) continue;
        const BLev &b = hc[c].lev[d];
        st.push_back({c, b.sgx * b.sgy * b.sgz});
        vc.push_back({c, b.nvb});
        rd.push_back({c, b.nrd});
        const int g = b.cs == 8 ? 0 : b.cs == 16 ? 1 : 2;
        ln[g].push_back({c, b.lgx});
      }
      add_tab(v < 0 ? B->st[d] : B->st_v[d][v], st);
      add_tab(v < 0 ? B->vc[d] : B->vc_v[d][v], vc);
      add_tab(v < 0 ? B->rd[d] : B->rd_v[d][v], rd);
      for (int g = 0; g < 3; g++) add_tab(v < 0 ? B->ln[d][g] : B->ln_v[d][v][g], ln[g]);
    }
  }
  for (const uint32_t e : tabs)
    if ((e & 0xffffffu) >= (1u << 24)) { cv_set_error("cv_batch_create: case too large"); b_release(B); return CV_ERR_ARG; }
  CV_CHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
  CV_CHECK(hipMalloc((void **)&B->dcases, sizeof(BCase) * ncases));
  CV_CHECK(hipMalloc((void **)&B->dctl, sizeof(BCtl)));
  CV_CHECK(hipMalloc((void **)&B->dbsc, sizeof(double) * BSC * ncases));
  CV_CHECK(hipMalloc((void **)&B->dgat, sizeof(double) * BSC * ncases));
  CV_CHECK(hipMalloc((void **)&B->dtab, sizeof(uint32_t) * std::max<size_t>(tabs.size(), 1)));
  CV_CHECK(hipHostMalloc((void **)&B->hctl, sizeof(BCtl), hipHostMallocDefault));
  CV_CHECK(hipHostMalloc((void **)&B->hbsc, sizeof(double) * BSC * ncases, hipHostMallocDefault));
  CV_CHECK(hipHostMalloc((void **)&B->hgat, sizeof(double) * BSC * ncases, hipHostMallocDefault));
  memset(B->hbsc, 0, sizeof(double) * BSC * ncases);
  for (int c = 0; c < ncases; c++) hc[c].bsc = B->dbsc + (long)c * BSC;
  CV_CHECK(hipMemcpy(B->dcases, hc.data(), sizeof(BCase) * ncases, hipMemcpyHostToDevice));
  if (!tabs.empty()) CV_CHECK(hipMemcpy(B->dtab, tabs.data(), sizeof(uint32_t) * tabs.size(), hipMemcpyHostToDevice));
  CV_CHECK(hipMemcpy(B->dbsc, B->hbsc, sizeof(double) * BSC * ncases, hipMemcpyHostToDevice));
  // the V-cycle from depth CV_BATCH_LEGS down as persistent legs ("off" or a
  // negative depth: one launch per operation throughout); the legs' line
  // items take up to 16 rows per thread
  {
    const char *el = getenv("CV_BATCH_LEGS");
    int d0 = el ? (strcmp(el, "off") == 0 ? -1 : atoi(el)) : CV_BATCH_LEGS_DEFAULT;
    for (int c = 0; c < ncases && d0 >= 0; c++)
      for (int d = 0; d < hc[c].nlev; d++)
        if (hc[c].lev[d].cs > 16) d0 = -1;
    if (d0 >= 0 && d0 < B->maxlev) {
      const char *ek = getenv("CV_BATCH_LEG_KC");
      const int kc = ek ? std::max(1, atoi(ek)) : 4;
      if ((rc = legs_build(B, hc, d0, kc, &B->legs))) { b_release(B); return rc; }
      if (hipHostMalloc((void **)&B->herr, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        b_release(B);
        cv_set_error("cv_batch_create: out of host memory");
        return CV_ERR_OOM;
      }
      *B->herr = 0;
    }
  }
  if ((rc = cv_mom_batch_create(cases, ncases, B->stream, &B->mom))) { b_release(B); return rc; }
  {
    std::vector<BDiv> hd(ncases);
    for (int c = 0; c < ncases; c++) {
      cv_ctx *x = cases[c];
      CvLevel &F = x->lev[x->nlev - 1];
      hd[c] = BDiv{idx_of(F), x->ucont[0], x->ucont[1], x->ucont[2], F.aj, F.nvert, F.d.n, x->part};
    }
    if (hipMalloc((void **)&B->ddiv, sizeof(BDiv) * ncases) != hipSuccess) {
      b_release(B);
      cv_set_error("cv_batch_create: out of device memory");
      return CV_ERR_OOM;
    }
    CV_CHECK(hipMemcpy(B->ddiv, hd.data(), sizeof(BDiv) * ncases, hipMemcpyHostToDevice));
  }
  B->gpre.assign(ncases, nullptr);
  B->gpost.assign(ncases, nullptr);
  B->ev_case.resize(ncases);
  for (int c = 0; c < ncases; c++) CV_CHECK(hipEventCreateWithFlags(&B->ev_case[c], hipEventDisableTiming));
  CV_CHECK(hipEventCreateWithFlags(&B->ev_solved, hipEventDisableTiming));
  CV_CHECK(hipEventCreate(&B->ev_t0));
  CV_CHECK(hipEventCreate(&B->ev_t1));
  *out = B;
  return CV_OK;
}

extern "C" int cv_batch_destroy(cv_batch *B) {
  if (!B) return CV_ERR_ARG;
  b_release(B);
  return CV_OK;
}

// every case must still have the options and Krylov space the batch was built on
static int b_check(cv_batch *B) {
  for (int c = 0; c < B->n; c++) {
    cv_ctx *x = B->cs[c];
    if (!same_opts(x->opt, B->opt) || x->krylov_m < B->opt.restart) {
      cv_set_error("cv_batch: solver options of case %d changed since cv_batch_create", c);
      return CV_ERR_STATE;
    }
    if (x->ib_gen != B->ib_gen[c]) {
      cv_set_error("cv_batch: IB nodes or nvert of case %d changed since cv_batch_create", c);
      return CV_ERR_STATE;
    }
  }
  return CV_OK;
}

// the batched solve on the batch stream once every case's stream has
// recorded ev_case; then each case's stream waits for the solve
static int b_solve(cv_batch *B, double *t_p) {
  for (int c = 0; c < B->n; c++) CV_CHECK(hipStreamWaitEvent(B->stream, B->ev_case[c], 0));
  CV_CHECK(hipEventRecord(B->ev_t0, B->stream));
  int rc = b_fgmres(B);
  if (rc) return rc;
  CV_CHECK(hipEventRecord(B->ev_t1, B->stream));
  CV_CHECK(hipEventRecord(B->ev_solved, B->stream));
  for (int c = 0; c < B->n; c++) CV_CHECK(hipStreamWaitEvent(B->cs[c]->stream, B->ev_solved, 0));
  if (t_p) {
    CV_CHECK(hipEventSynchronize(B->ev_t1));
    float ms = 0;
    CV_CHECK(hipEventElapsedTime(&ms, B->ev_t0, B->ev_t1));
    *t_p = ms * 1e-3;
  }
  return CV_OK;
}

// VolumeFlux_rev and PoissonRHS of a case (statistics stay on the device)
static int b_rhs_work(cv_ctx *x) {
  int rc = cv_volumeflux_rev_dev(x, nullptr);
  return rc ? rc : cv_poisson_rhs_dev(x, nullptr);
}

// a case's kernels of `work` captured once on its stream into *g, then
// replayed; the first call captures and launches
template <class W>
static int b_graph_run(cv_ctx *x, hipGraphExec_t *g, W &&work) {
  if (!*g) {
    hipGraph_t gg = nullptr;
    CV_CHECK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeRelaxed));
    const int rc = work();
    const hipError_t e = hipStreamEndCapture(x->stream, &gg);
    if (rc) {
      if (e == hipSuccess && gg) (void)hipGraphDestroy(gg);
      return rc;
    }
    if (e != hipSuccess) { cv_set_error("hipStreamEndCapture: %s", hipGetErrorString(e)); return CV_ERR_HIP; }
    const hipError_t ei = hipGraphInstantiate(g, gg, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gg);
    if (ei != hipSuccess) { *g = nullptr; cv_set_error("hipGraphInstantiate: %s", hipGetErrorString(ei)); return CV_ERR_HIP; }
  }
  CV_CHECK(hipGraphLaunch(*g, x->stream));
  return CV_OK;
}

// the per-case work before the solve is a fixed kernel sequence once the
// BDF history exists (steps_done > 0) for explicit RK with no kernel timing;
// ImpRK synchronises with the host and always runs eagerly
static bool b_pre_graphable(const cv_batch *B, const cv_ctx *x) {
  return B->use_graph && x->steps_done > 0 && x->mom.scheme == CV_MOM_RK && x->prof.kind == CV_PROF_NONE;
}

// statistics of the solve (and of the step) of every case, read back after
// the case's stream has finished
static int b_stats(cv_batch *B, cv_step_stats *st, double t_p) {
  for (int c = 0; c < B->n; c++) {
    cv_ctx *x = B->cs[c];
    CV_CHECK(hipMemcpyAsync(x->red_host + 100, x->dsc + 104, sizeof(double), hipMemcpyDeviceToHost, x->stream));
    CV_CHECK(hipMemcpyAsync(x->red_host + 104, x->dsc + 112, 4 * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  }
  for (int c = 0; c < B->n; c++) {
    cv_ctx *x = B->cs[c];
    CV_CHECK(hipStreamSynchronize(x->stream));
    cv_step_stats &s = st[c];
    const BFg &f = B->fg[c];
    s.ksp_its = f.its; s.ksp_rnorm = f.rnorm; s.ksp_bnorm = f.bnorm;
    s.rhs_sum = x->red_host[100];
    s.ibm_flux0 = x->red_host[104]; s.ibm_area0 = x->red_host[105];
    s.ibm_flux1 = x->red_host[106]; s.ibm_area1 = x->red_host[107];
    s.converged = B->rc[c] == CV_OK;
    s.t_p = t_p;
  }
  return CV_OK;
}

extern "C" int cv_batch_poisson_solve(cv_batch *B, cv_step_stats *st) {
  if (!B || !st) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(B->device));
  int rc = b_check(B);
  if (rc) return rc;
  for (int c = 0; c < B->n; c++) st[c] = cv_step_stats{};
  for (int c = 0; c < B->n; c++) {
    if ((rc = b_rhs_work(B->cs[c]))) return rc;
    CV_CHECK(hipEventRecord(B->ev_case[c], B->cs[c]->stream));
  }
  double tp = 0;
  if ((rc = b_solve(B, &tp))) return rc;
  if ((rc = b_stats(B, st, tp))) return rc;
  int worst = CV_OK;
  for (int c = 0; c < B->n; c++)
    if (B->rc[c] == CV_ERR_NONFINITE || (B->rc[c] && !worst)) worst = B->rc[c];
  return worst;
}

static int b_finish(cv_batch *B, cv_step_stats *st, double tp, bool batched = false);

// every case runs explicit RK with kernel timing off and has only wall and
// periodic faces: the whole step around the solve can then run on the
// batch stream, one launch per kernel (IB reconstruction per IB case)
static bool b_mom_batchable(const cv_batch *B) {
  if (!B->use_mom || !B->mom) return false;
  for (const cv_ctx *x : B->cs)
    if (x->mom.scheme != CV_MOM_RK || x->prof.kind != CV_PROF_NONE || x->ubf) return false;
  return true;
}

#define B2_LAUNCH(kern, t, pm, ...)                                                                          \
  do {                                                                                                      \
    if ((t).len > 0 && (pm))                                                                                \
      hipLaunchKernelGGL(kern, dim3((unsigned)(t).len), dim3(BS), 0, B->stream, (const BCase *)B->dcases,   \
                         (const uint32_t *)BTAB(t), (unsigned long long)(pm), ##__VA_ARGS__);               \
  } while (0)

// rotation + RK (cv_step_pre_dev), VolumeFlux_rev and PoissonRHS of every
// case on the batch stream
static int b_pre_batched(cv_batch *B) {
  unsigned long long rot = 0;
  for (int c = 0; c < B->n; c++) {   // the batch stream follows each case's earlier work
    CV_CHECK(hipEventRecord(B->ev_case[c], B->cs[c]->stream));
    CV_CHECK(hipStreamWaitEvent(B->stream, B->ev_case[c], 0));
    if (B->cs[c]->steps_done > 0) rot |= bit(c);
  }
  int rc = cv_mom_batch_rk(B->mom, rot);
  if (rc) return rc;
  const unsigned long long all = B->all;
  const unsigned nc = (unsigned)B->n;
  for (int pass = 0; pass < 3; pass++) {
    const int mode = pass == 1 ? 1 : 0;
    B2_LAUNCH(kb_ibm_flux, B->rd[0], all, mode);
    if (!mode) {
      hipLaunchKernelGGL(kb_final2, dim3(nc, 2), dim3(BS), 0, B->stream, (const BCase *)B->dcases, all,
                         pass == 0 ? 112 : 114);
      if (pass == 0) hipLaunchKernelGGL(kb_ibm_corr, dim3(nc), dim3(64), 0, B->stream, (const BCase *)B->dcases, all);
    }
  }
  B2_LAUNCH(kb_poisson_rhs, B->rd[0], all);
  hipLaunchKernelGGL(kb_final2, dim3(nc, 1), dim3(BS), 0, B->stream, (const BCase *)B->dcases, all, 104);
  CV_CHECK(hipGetLastError());
  CV_CHECK(hipEventRecord(B->ev_t0, B->stream));
  rc = b_fgmres(B);
  if (rc) return rc;
  CV_CHECK(hipEventRecord(B->ev_t1, B->stream));
  return CV_OK;
}

// UpdatePressure and Projection (with Contra2Cart and the ghost update) of
// every case on the batch stream; each case's stream then follows
static int b_post_batched(cv_batch *B) {
  const unsigned long long all = B->all;
  B2_LAUNCH(kb_p_add, B->vc[0], all);
  B2_LAUNCH(kb_pplane, B->vc[0], B->xper, 0, 0);
  B2_LAUNCH(kb_pplane, B->vc[0], B->zper, 2, 0);
  B2_LAUNCH(kb_pplane, B->vc[0], B->xper, 0, 1);
  B2_LAUNCH(kb_pplane, B->vc[0], B->zper, 2, 1);
  B2_LAUNCH(kb_projection, B->vc[0], all);
  int rc = cv_mom_batch_ghost(B->mom, 0, 1);
  if (rc) return rc;
  CV_CHECK(hipEventRecord(B->ev_solved, B->stream));
  for (int c = 0; c < B->n; c++) CV_CHECK(hipStreamWaitEvent(B->cs[c]->stream, B->ev_solved, 0));
  return CV_OK;
}

extern "C" int cv_batch_step(cv_batch *B, cv_step_stats *st) {
  if (!B || !st) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(B->device));
  int rc = b_check(B);
  if (rc) return rc;
  for (int c = 0; c < B->n; c++) st[c] = cv_step_stats{};
  double tp = 0;
  if (b_mom_batchable(B)) {
    if ((rc = b_pre_batched(B))) return rc;
    if ((rc = b_post_batched(B))) return rc;
    CV_CHECK(hipEventSynchronize(B->ev_t1));
    float ms = 0;
    CV_CHECK(hipEventElapsedTime(&ms, B->ev_t0, B->ev_t1));
    tp = ms * 1e-3;
    return b_finish(B, st, tp, true);
  }
  // momentum, VolumeFlux_rev and PoissonRHS per case on its own stream
  for (int c = 0; c < B->n; c++) {
    cv_ctx *x = B->cs[c];
    cv_step_stats *sc = &st[c];
    auto pre = [x, sc]() {
      const int r = cv_step_pre_dev(x, sc);
      return r ? r : b_rhs_work(x);
    };
    rc = b_pre_graphable(B, x) ? b_graph_run(x, &B->gpre[c], pre) : pre();
    if (rc) return rc;
    CV_CHECK(hipEventRecord(B->ev_case[c], x->stream));
  }
  if ((rc = b_solve(B, &tp))) return rc;
  // UpdatePressure and Projection per case
  for (int c = 0; c < B->n; c++) {
    cv_ctx *x = B->cs[c];
    auto post = [x]() { return cv_step_post_dev(x); };
    rc = (B->use_graph && x->prof.kind == CV_PROF_NONE) ? b_graph_run(x, &B->gpost[c], post) : post();
    if (rc) return rc;
  }
  return b_finish(B, st, tp);
}

// statistics, Divergence and the step count of every case
// batched (the whole step ran on the batch stream): Divergence of every
// case in two launches and one read, then the IB reconstruction after it
// (solvers.c:1095-1101) in one batched sweep, and the case streams wait for
// it; otherwise both per case on the case's stream
static int b_finish(cv_batch *B, cv_step_stats *st, double tp, bool batched) {
  int rc;
  if ((rc = b_stats(B, st, tp))) return rc;
  if (batched) {
    hipLaunchKernelGGL(kbd_divmax, dim3(CV_RED_GRID, (unsigned)B->n), dim3(BS), 0, B->stream, (const BDiv *)B->ddiv,
                       CV_RED_GRID);
    hipLaunchKernelGGL(kbd_final, dim3((unsigned)B->n), dim3(BS), 0, B->stream, (const BDiv *)B->ddiv, CV_RED_GRID,
                       B->dgat);
    CV_CHECK(hipMemcpyAsync(B->hgat, B->dgat, 2 * (size_t)B->n * sizeof(double), hipMemcpyDeviceToHost, B->stream));
    CV_CHECK(hipStreamSynchronize(B->stream));
    for (int c = 0; c < B->n; c++) { st[c].maxdiv = B->hgat[2 * c]; st[c].maxdiv_idx = (long)B->hgat[2 * c + 1]; }
  }
  int conv_rc = CV_OK;
  for (int c = 0; c < B->n; c++) {
    cv_ctx *x = B->cs[c];
    if (x->inlet || x->outlet) {
      CV_CHECK(hipMemcpyAsync(x->red_host + 110, x->dsc + CV_DSC_FLUXIN, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              x->stream));
      if (batched) CV_CHECK(hipStreamSynchronize(x->stream));
    }
    if (!batched && (rc = cv_divergence_dev(x, &st[c].maxdiv, &st[c].maxdiv_idx))) return rc;
    if (!batched && (rc = cv_ib_dev(x))) return rc;     // solvers.c:1095-1101
    if (x->inlet || x->outlet) { st[c].flux_in = x->red_host[110]; st[c].flux_out = x->red_host[111]; }
    x->last = st[c];
    x->steps_done++;
    // every case's record is complete before a failure is reported; a
    // non-finite case outranks an unconverged one
    const int rcc = !std::isfinite(st[c].maxdiv) ? CV_ERR_NONFINITE : B->rc[c];
    if (rcc == CV_ERR_NONFINITE) { st[c].converged = 0; conv_rc = CV_ERR_NONFINITE; }
    else if (rcc && !conv_rc) conv_rc = rcc;
  }
  if (batched) {
    if ((rc = cv_mom_batch_ib(B->mom))) return rc;
    CV_CHECK(hipEventRecord(B->ev_solved, B->stream));
    for (int c = 0; c < B->n; c++) CV_CHECK(hipStreamWaitEvent(B->cs[c]->stream, B->ev_solved, 0));
  }
  return conv_rc;
}
