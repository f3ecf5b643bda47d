// cv_internal.h -- private declarations of the CURVIB fractional-step
// library for gfx950 (MI355X).  Device layout: every field is planar
// (structure of arrays) in DMDA storage order, index i + mx*(j + my*k),
// i fastest; a Cmpnts field is three planar arrays x|y|z.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <cstdio>
#include <vector>
#include <string>
#include "../../include/curvib_hip.h"

#define CV_MAXLEV 10

#define CV_CHECK(expr)                                                          \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      cv_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
      return CV_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

void cv_set_error(const char *fmt, ...);

// per-cell operator word (see cv_poisson.hip):
//   bit 0           fluid row
//   bits 1+5d       face d (upper face of the cell) enabled
//   bits 2+5d..3+5d cross mode of the first tangential axis of face d
//   bits 4+5d..5+5d cross mode of the second tangential axis of face d
// tangential axes of face d: d=0 -> (1,2), d=1 -> (0,2), d=2 -> (0,1)
enum { XM_CENTRAL = 0, XM_LOWER = 1, XM_UPPER = 2, XM_DROP = 3 };

struct CvDims {
  int mx, my, mz;
  long n;
  int xper, zper;
  int isc, jsc, ksc;
};

struct CvLevel {
  CvDims d;
  double *nvert;          // storage nvert (level nvert hierarchy)
  float *nwf;             // max(0, 1 - nvert) in FP32: restriction weights of the FP32 V-cycle
  double *aj;             // cell 1/J
  double *g[9];           // g11 g12 g13 (i faces) g21 g22 g23 (j faces) g31 g32 g33 (k faces)
  uint32_t *flags;
  double *dinv;           // 1/diag of the operator row (smoother)
  float *gf[9], *ajf, *dinvf;  // FP32 copies for the mixed-precision V-cycle
  void *gh;               // CV_MIXED16 packs (null until first used): n x 16 bytes (g0..g7 *
                          // 2^hexp in FP16), n x 8 bytes (g8 * 2^hexp in FP16, unused FP16,
                          // aj * 2^-hexp in FP32), n x 8 bytes (y-line a, ip * 2^yexp, cp in FP16)
  int hexp, yexp;
  double *ylo, *yip, *ycp;     // y-line smoother: lower coefficient, Thomas 1/pivot and c'
  float *ylof, *yipf, *ycpf;
  double *ltmp; float *ltmpf;  // line-solve forward-sweep scratch
  double *x, *b, *r;      // FP64 V-cycle work vectors
  float *xf, *bf, *rf;    // FP32 V-cycle work vectors
  double *coor;           // node coordinates, planar x|y|z (storage grid)
};

struct CvMetrics {        // fine-level metrics for the momentum right-hand side
  double *csi[3], *eta[3], *zet[3], *aj;
  double *icsi[3], *ieta[3], *izet[3], *iaj;
  double *jcsi[3], *jeta[3], *jzet[3], *jaj;
  double *kcsi[3], *keta[3], *kzet[3], *kaj;
};

// per-launch device-event timing of one selected kernel class
struct CvProf {
  int kind = CV_PROF_NONE, level = -1, prec = -1;
  std::vector<hipEvent_t> ev;   // start/end pairs of pending launches
  int pending = 0;
  long launches = 0;
  double ms = 0.0, bytes = 0.0;
};

struct cv_ctx {
  int device;
  hipStream_t stream;
  int nlev;
  int IM, JM, KM;
  int bctype[6];
  double ren, dt, st;
  cv_solver_opts opt;
  CvLevel lev[CV_MAXLEV];
  CvMetrics m;
  // fine-level state (planar)
  double *ucont[3], *ucont_o[3], *ucat[3], *rhs[3], *p, *phi, *bvec;
  double *ucont_rm1[3], *pucont[3];   // ImpRK: U^{n-1} and the pseudo-time base
  int steps_done;                      // steps since Ucont was last loaded (BDF1 on the first)
  cv_momentum_opts mom;
  double *rct[3];
  double *stage;          // 3n staging of the host field exchange (Cmpnts AoS)
  // Krylov space
  double *V, *Z, *w;      // (restart+1)*n, restart*n, n
  double *red;            // reduction scratch (device)
  double *red_host;       // pinned host mirror
  int red_len;
  cv_step_stats last;
  double nfluid[CV_MAXLEV];   // fluid-row count per level (nullspace mean)
  double *dsc;                // device scalars (64)
  double *part;               // reduction partials (33 * CV_RED_GRID)
  int krylov_m;               // restart length the Krylov space was sized for
  size_t bytes_alloc;
  std::vector<void *> allocs;
  int geometry_set, nvert_set;
  CvProf prof;
  // boundary faces other than walls and periodic pairs: inlet (bctype[4]
  // == 5), outlet (bctype[5] == 4), slip (3) and far field (6) faces
  int inlet, outlet;
  int slip, far;              // face masks, bit f for face f (0 i-, 1 i+, 2 j-, 3 j+, 4 k-, 5 k+)
  int inletprofile;           // -inlet
  double cm_c[3];             // -x_c -y_c -z_c
  double *ubf;                // Ubcs of the six faces (null with walls and periodic faces only)
  // immersed-boundary nodes (cv_set_ib_nodes) and the face and cell lists
  // of ibm_interpolation_advanced, built from nvert
  int ib_n, ib_mode;
  int *ib_node;               // per node: storage index of the node and of its three stencil cells
  double *ib_w;               // per node: cr1 cr2 cr3, d_s, d_i, ua xyz, dun, nf xyz (12)
  double *ib_new;             // per node: the reconstructed Ucat xyz and p (4)
  int ib_nface[3];            // faces touching an IB node, per direction
  int *ib_face;               // their cells (upper face), x then y then z
  int ib_ncell;
  int *ib_cell;               // the cells those faces bound (Contra2Cart)
  int ib_gen;                 // bumped whenever the IB buffers are replaced (captured graphs hold them)
  hipEvent_t step_ev[4];      // cv_flow_step's phase timing (momentum, solve, projection)
};

// device scalars of the inlet/outlet update (slots of cv_ctx::dsc)
#define CV_DSC_FLUXIN  116    // FluxInSum (InflowFlux)
#define CV_DSC_FLUXOUT 117    // FluxOutSum after FormBCS's outlet correction
#define CV_DSC_OUTSUM  118    // outlet velocity flux and area over the last fluid cells (2)
#define CV_DSC_OUTDIAG 121    // OutflowFlux's sum
#define CV_DSC_FAR     122    // far-field inflow, outflow, |flux|, inflow area, outflow area (5)

// Records device events around one launch when (kind, level, precision)
// matches the selection; prec 0 = FP64, 1 = FP32.  bytes: algorithmic bytes
// the launch moves.
struct CvProfScope {
  cv_ctx *c;
  bool on;
  CvProfScope(cv_ctx *c, int kind, int level, int prec, double bytes);
  ~CvProfScope();
};
int cv_prof_drain(cv_ctx *c);

// helpers implemented in cv_ctx.hip
int cv_alloc(cv_ctx *c, void **p, size_t bytes);
void cv_free(cv_ctx *c, void *p);
int cv_reduce_sum(cv_ctx *c, const double *const *vecs_a, const double *const *vecs_b, int nvec,
                  long n, double *out_host);

static inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// kernels shared across translation units
__global__ void k_copy(double *dst, const double *src, long n);
__global__ void k_fill(double *dst, double v, long n);
__global__ void k_axpby(double *y, double a, const double *x, double b, long n);   // y = a*x + b*y

int cv_metrics_build(cv_ctx *c, const double *xyz_host);
int cv_levels_setup_nvert(cv_ctx *c, const double *nvert_host);
long cv_half_bytes(const cv_grid_desc *desc);
// cv_set_geometry's transient device scratch: 17 padded Cmpnts fields and 4
// padded scalars of the finest level, plus its node coordinates
#define CV_METRIC_SCRATCH_VEC 17
#define CV_METRIC_SCRATCH_SC 4
long cv_geometry_scratch_bytes(const cv_grid_desc *desc);
int cv_poisson_prepare(cv_ctx *c);
int cv_contra2cart_dev(cv_ctx *c);
int cv_ghost_update(cv_ctx *c, int formbcs, bool cart = true);
int cv_inflow_dev(cv_ctx *c);
int cv_bc_update_dev(cv_ctx *c);   // InflowFlux + FormBCS
int cv_ib_dev(cv_ctx *c);          // ibm_interpolation_advanced (nothing without IB nodes)
int cv_ib_dev_on(cv_ctx *c, hipStream_t s);   // the same on another stream
void cv_ib_clear(cv_ctx *c);
int cv_poisson_solve_dev(cv_ctx *c, cv_step_stats *st);
int cv_update_pressure_dev(cv_ctx *c);
int cv_projection_dev(cv_ctx *c);
int cv_divergence_dev(cv_ctx *c, double *maxdiv, long *idx);
int cv_volumeflux_rev_dev(cv_ctx *c, double *fa4);
int cv_poisson_rhs_dev(cv_ctx *c, double *rhs_sum);
int cv_ff1_dev(cv_ctx *c);
int cv_rk_dev(cv_ctx *c);
int cv_imprk_dev(cv_ctx *c, cv_step_stats *st);
int cv_calcrhs_dev(cv_ctx *c);
int cv_step_pre_dev(cv_ctx *c, cv_step_stats *st);
int cv_step_post_dev(cv_ctx *c);

// batched momentum over many contexts on one stream (cv_momentum.hip)
struct CvMomBatch;
int cv_mom_batch_create(cv_ctx *const *cases, int n, hipStream_t s, CvMomBatch **out);
void cv_mom_batch_destroy(CvMomBatch *mb);
int cv_mom_batch_rk(CvMomBatch *mb, unsigned long long rot);
int cv_mom_batch_ghost(CvMomBatch *mb, int formbcs, int cart);
int cv_mom_batch_ib(CvMomBatch *mb);   // the IB reconstruction of every case with IB nodes, on the batch stream

__device__ inline int cv_wrap(int i, int m, int per) {
  return per ? (i == m - 1 ? 1 : (i == 0 ? m - 2 : i)) : i;
}

// line-solve launches of at most this many blocks (one per CU of the
// MI355X's 256) load x with the residual (yline_reg_tile's PX)
#define CV_PX_BLOCKS 256

// XCD-aware tile order.  Work-groups are handed to the 8 XCDs of the chip
// round-robin by linear id; this maps the blocks one XCD receives onto a
// contiguous run of tiles (x fastest), so tiles that share halo lines share
// that XCD's L2.  A bijection on [0, gridDim.x*gridDim.y*gridDim.z).
#define CV_NXCD 8
// Interleaved runs: tile t belongs to run t / R; run r goes to XCD r mod 8,
// so every XCD holds runs of R consecutive tiles (which share halo lines in
// its L2) spread evenly over the whole grid.  Blocks past the last full
// round of 8 runs keep their own index.  A bijection on the block range.
__device__ inline void cv_xcd_runs(int R, int &tx, int &ty, int &tz) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned full = (nb / (CV_NXCD * (unsigned)R)) * CV_NXCD * (unsigned)R;
  unsigned t = b;
  if (b < full) {
    const unsigned x = b % CV_NXCD, l = b / CV_NXCD;
    t = (x + CV_NXCD * (l / (unsigned)R)) * (unsigned)R + l % (unsigned)R;
  }
  tx = (int)(t % gx);
  ty = (int)((t / gx) % gy);
  tz = (int)(t / (gx * gy));
}
__device__ inline void cv_xcd_tile(int &tx, int &ty, int &tz) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned q = nb / CV_NXCD, rem = nb % CV_NXCD;
  const unsigned x = b % CV_NXCD, l = b / CV_NXCD;
  const unsigned t = (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + l;
  tx = (int)(t % gx);
  ty = (int)((t / gx) % gy);
  tz = (int)(t / (gx * gy));
}

