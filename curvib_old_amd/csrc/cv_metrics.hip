// cv_metrics.hip -- FP64 curvilinear metrics on the device: FormMetrics
// (metrics.c:4-1229) followed by Update_Metrics_PBC
// (implicitsolver.c:2174-2573), for every multigrid level, plus the
// coordinate restriction of MG_Initial (init.c:786-828).
//
// The computation runs on temporary padded arrays (3 ghost layers per side,
// the local-vector view of a single-rank DMDA with stencil width 3) so that
// the index arithmetic of the face-centre ghost construction and of the
// periodic metric fix-up is the reference's own; the storage region is then
// copied into the compact planar arrays used by the hot kernels.
#include "cv_internal.h"

namespace {

constexpr int PG = 3;

struct PadGeom {
  int mx, my, mz, px, py, pz;
  int xper, zper;
  __host__ __device__ long ix(int i, int j, int k) const {
    return (long)(i + PG) + (long)px * ((long)(j + PG) + (long)py * (k + PG));
  }
  __host__ __device__ long np() const { return (long)px * py * pz; }
};

struct PV { double *x, *y, *z; };

template <class F>
__global__ void k_box(int i0, int i1, int j0, int j1, int k0, int k1, F f) {
  const long ni = i1 - i0, nj = j1 - j0, nk = k1 - k0;
  const long n = ni * nj * nk;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i = i0 + (int)(t % ni);
    int j = j0 + (int)((t / ni) % nj);
    int k = k0 + (int)(t / (ni * nj));
    f(i, j, k);
  }
}

template <class F>
void box(hipStream_t s, int i0, int i1, int j0, int j1, int k0, int k1, F f) {
  long n = (long)(i1 - i0) * (j1 - j0) * (k1 - k0);
  if (n <= 0) return;
  long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_box<F>, dim3((unsigned)nb), dim3(256), 0, s, i0, i1, j0, j1, k0, k1, f);
}

// DMGlobalToLocal of a single rank: periodic ghosts take the value of
// global point (l mod m); non-periodic ghosts are left alone.
void g2l(hipStream_t s, const PadGeom g, double *a) {
  box(s, -PG, g.mx + PG, -PG, g.my + PG, -PG, g.mz + PG, [=] __device__(int i, int j, int k) {
    bool iin = i >= 0 && i < g.mx, jin = j >= 0 && j < g.my, kin = k >= 0 && k < g.mz;
    if (iin && jin && kin) return;
    if (!jin) return;                 // y is never periodic
    if (!iin && !g.xper) return;
    if (!kin && !g.zper) return;
    int ii = ((i % g.mx) + g.mx) % g.mx, kk = ((k % g.mz) + g.mz) % g.mz;
    a[g.ix(i, j, k)] = a[g.ix(ii, j, kk)];
  });
}

__device__ inline void cross3(double ax, double ay, double az, double bx, double by, double bz,
                              double *o0, double *o1, double *o2) {
  *o0 = ay * bz - az * by;
  *o1 = -ax * bz + az * bx;
  *o2 = ax * by - ay * bx;
}

__device__ inline double det3(double dxdc, double dydc, double dzdc, double dxde, double dyde,
                              double dzde, double dxdz, double dydz, double dzdz) {
  return dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
         dzdc * (dxde * dydz - dyde * dxdz);
}

struct MetricSet {       // padded arrays of one level
  PV coor, csi, eta, zet, icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, gs, cx, cy, cz;
  double *aj, *iaj, *jaj, *kaj;
};

}  // namespace

// face-centred metrics of one face family: shared by the I/J/K passes.
// kind 0/1/2 selects which difference rules apply (metrics.c:713-790,
// 872-942, 1019-1090).
static int level_metrics(cv_ctx *c, int l, const double *fine_xyz_host) {
  CvLevel &L = c->lev[l];
  hipStream_t s = c->stream;
  const int mx = L.d.mx, my = L.d.my, mz = L.d.mz;
  const int IM = mx - 1, JM = my - 1, KM = mz - 1;
  PadGeom g{mx, my, mz, mx + 2 * PG, my + 2 * PG, mz + 2 * PG, L.d.xper, L.d.zper};
  const long np = g.np();
  const int xper = L.d.xper, zper = L.d.zper;
  const int bct0 = c->bctype[0], bct4 = c->bctype[4];
  // padded temporaries, freed on every return path (an early return waits
  // for the stream first, so no queued kernel still uses them)
  struct Scratch {
    double *p = nullptr;
    hipStream_t s;
    ~Scratch() {
      if (p) { (void)hipStreamSynchronize(s); (void)hipFree(p); }
    }
  };
  const int nvec = CV_METRIC_SCRATCH_VEC, nsc = CV_METRIC_SCRATCH_SC;
  Scratch sbuf{nullptr, s}, sxyz{nullptr, s};
  CV_CHECK(hipMalloc((void **)&sbuf.p, (size_t)(3 * nvec + nsc) * np * sizeof(double)));
  double *buf = sbuf.p;
  CV_CHECK(hipMemsetAsync(buf, 0, (size_t)(3 * nvec + nsc) * np * sizeof(double), s));
  MetricSet M;
  PV *pv[] = {&M.coor, &M.csi, &M.eta, &M.zet, &M.icsi, &M.ieta, &M.izet, &M.jcsi, &M.jeta, &M.jzet,
              &M.kcsi, &M.keta, &M.kzet, &M.gs, &M.cx, &M.cy, &M.cz};
  for (int q = 0; q < nvec; q++) {
    pv[q]->x = buf + (3 * q + 0) * np;
    pv[q]->y = buf + (3 * q + 1) * np;
    pv[q]->z = buf + (3 * q + 2) * np;
  }
  M.aj = buf + (3 * nvec + 0) * np; M.iaj = buf + (3 * nvec + 1) * np;
  M.jaj = buf + (3 * nvec + 2) * np; M.kaj = buf + (3 * nvec + 3) * np;

  // ---- node coordinates (storage index IM keeps 0, init.c:666) ----
  PV co = M.coor;
  if (l == c->nlev - 1) {
    long nn = (long)IM * JM * KM;
    CV_CHECK(hipMalloc((void **)&sxyz.p, 3 * nn * sizeof(double)));
    const double *d_xyz = sxyz.p;
    CV_CHECK(hipMemcpyAsync(sxyz.p, fine_xyz_host, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
    box(s, 0, IM, 0, JM, 0, KM, [=] __device__(int i, int j, int k) {
      long n = (long)i + (long)IM * (j + (long)JM * k);
      long p = g.ix(i, j, k);
      co.x[p] = d_xyz[3 * n]; co.y[p] = d_xyz[3 * n + 1]; co.z[p] = d_xyz[3 * n + 2];
    });
    CV_CHECK(hipStreamSynchronize(s));
    CV_CHECK(hipFree(sxyz.p));
    sxyz.p = nullptr;
  } else {
    // coarse coordinates = fine coordinates at (2i,2j,2k) (GridRestriction)
    const CvLevel &H = c->lev[l + 1];
    const int hmx = H.d.mx, hmy = H.d.my;
    const double *hc = H.coor;
    const long hn = H.d.n;
    const int isc = L.d.isc, jsc = L.d.jsc, ksc = L.d.ksc;
    box(s, 0, mx - 1, 0, my - 1, 0, mz - 1, [=] __device__(int i, int j, int k) {
      int ih = isc ? i : 2 * i, jh = jsc ? j : 2 * j, kh = ksc ? k : 2 * k;
      long q = (long)ih + (long)hmx * (jh + (long)hmy * kh);
      long p = g.ix(i, j, k);
      co.x[p] = hc[q]; co.y[p] = hc[hn + q]; co.z[p] = hc[2 * hn + q];
    });
  }
  // keep a compact copy of this level's coordinates for the next coarser one
  {
    double *dst = L.coor;
    const long n = L.d.n;
    box(s, 0, mx, 0, my, 0, mz, [=] __device__(int i, int j, int k) {
      long q = (long)i + (long)mx * (j + (long)my * k), p = g.ix(i, j, k);
      dst[q] = co.x[p]; dst[n + q] = co.y[p]; dst[2 * n + q] = co.z[p];
    });
  }
  g2l(s, g, co.x); g2l(s, g, co.y); g2l(s, g, co.z);

  // ---- csi, eta, zet, aj (metrics.c:80-212) ----
  PV csi = M.csi, eta = M.eta, zet = M.zet;
  double *aj = M.aj;
  box(s, 0, mx - 1, 0, my - 1, 0, mz - 1, [=] __device__(int i, int j, int k) {
#define X(a, b, cc) co.x[g.ix(cc, b, a)]
#define Y(a, b, cc) co.y[g.ix(cc, b, a)]
#define Z(a, b, cc) co.z[g.ix(cc, b, a)]
    long p = g.ix(i, j, k);
    double dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz;
    if (k >= 1 && j >= 1) {   // csi: k in [1,mz-1), j in [1,my-1), i in [0,mx-1)
      dxde = 0.5 * (X(k, j, i) + X(k - 1, j, i) - X(k, j - 1, i) - X(k - 1, j - 1, i));
      dyde = 0.5 * (Y(k, j, i) + Y(k - 1, j, i) - Y(k, j - 1, i) - Y(k - 1, j - 1, i));
      dzde = 0.5 * (Z(k, j, i) + Z(k - 1, j, i) - Z(k, j - 1, i) - Z(k - 1, j - 1, i));
      dxdz = 0.5 * (X(k, j - 1, i) + X(k, j, i) - X(k - 1, j - 1, i) - X(k - 1, j, i));
      dydz = 0.5 * (Y(k, j - 1, i) + Y(k, j, i) - Y(k - 1, j - 1, i) - Y(k - 1, j, i));
      dzdz = 0.5 * (Z(k, j - 1, i) + Z(k, j, i) - Z(k - 1, j - 1, i) - Z(k - 1, j, i));
      cross3(dxde, dyde, dzde, dxdz, dydz, dzdz, &csi.x[p], &csi.y[p], &csi.z[p]);
    }
    if (k >= 1 && j >= 1 && i >= 1) {   // eta
      dxdc = 0.5 * (X(k, j, i) + X(k - 1, j, i) - X(k, j, i - 1) - X(k - 1, j, i - 1));
      dydc = 0.5 * (Y(k, j, i) + Y(k - 1, j, i) - Y(k, j, i - 1) - Y(k - 1, j, i - 1));
      dzdc = 0.5 * (Z(k, j, i) + Z(k - 1, j, i) - Z(k, j, i - 1) - Z(k - 1, j, i - 1));
      dxdz = 0.5 * (X(k, j, i) + X(k, j, i - 1) - X(k - 1, j, i) - X(k - 1, j, i - 1));
      dydz = 0.5 * (Y(k, j, i) + Y(k, j, i - 1) - Y(k - 1, j, i) - Y(k - 1, j, i - 1));
      dzdz = 0.5 * (Z(k, j, i) + Z(k, j, i - 1) - Z(k - 1, j, i) - Z(k - 1, j, i - 1));
      cross3(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &eta.x[p], &eta.y[p], &eta.z[p]);
    }
    if (j >= 1 && i >= 1) {   // zet: k from 0
      dxdc = 0.5 * (X(k, j, i) + X(k, j - 1, i) - X(k, j, i - 1) - X(k, j - 1, i - 1));
      dydc = 0.5 * (Y(k, j, i) + Y(k, j - 1, i) - Y(k, j, i - 1) - Y(k, j - 1, i - 1));
      dzdc = 0.5 * (Z(k, j, i) + Z(k, j - 1, i) - Z(k, j, i - 1) - Z(k, j - 1, i - 1));
      dxde = 0.5 * (X(k, j, i) + X(k, j, i - 1) - X(k, j - 1, i) - X(k, j - 1, i - 1));
      dyde = 0.5 * (Y(k, j, i) + Y(k, j, i - 1) - Y(k, j - 1, i) - Y(k, j - 1, i - 1));
      dzde = 0.5 * (Z(k, j, i) + Z(k, j, i - 1) - Z(k, j - 1, i) - Z(k, j - 1, i - 1));
      cross3(dxdc, dydc, dzdc, dxde, dyde, dzde, &zet.x[p], &zet.y[p], &zet.z[p]);
    }
    if (k >= 1 && j >= 1 && i >= 1) {   // Jacobian
      // a=(k,j,i) b=(k,j-1,i) c=(k-1,j,i) d=(k-1,j-1,i), e..h the same at i-1
      double ax = X(k, j, i), bx = X(k, j - 1, i), cx = X(k - 1, j, i), dx = X(k - 1, j - 1, i);
      double ex = X(k, j, i - 1), fx = X(k, j - 1, i - 1), gx = X(k - 1, j, i - 1), hx = X(k - 1, j - 1, i - 1);
      double ay = Y(k, j, i), by = Y(k, j - 1, i), cy = Y(k - 1, j, i), dy = Y(k - 1, j - 1, i);
      double ey = Y(k, j, i - 1), fy = Y(k, j - 1, i - 1), gy = Y(k - 1, j, i - 1), hy = Y(k - 1, j - 1, i - 1);
      double az = Z(k, j, i), bz = Z(k, j - 1, i), cz = Z(k - 1, j, i), dz = Z(k - 1, j - 1, i);
      double ez = Z(k, j, i - 1), fz = Z(k, j - 1, i - 1), gz = Z(k - 1, j, i - 1), hz = Z(k - 1, j - 1, i - 1);
      dxdc = 0.25 * (ax + bx + cx + dx - ex - fx - gx - hx);
      dydc = 0.25 * (ay + by + cy + dy - ey - fy - gy - hy);
      dzdc = 0.25 * (az + bz + cz + dz - ez - fz - gz - hz);
      dxde = 0.25 * (ax + ex + cx + gx - bx - fx - dx - hx);
      dyde = 0.25 * (ay + ey + cy + gy - by - fy - dy - hy);
      dzde = 0.25 * (az + ez + cz + gz - bz - fz - dz - hz);
      dxdz = 0.25 * (ax + bx + ex + fx - cx - dx - gx - hx);
      dydz = 0.25 * (ay + by + ey + fy - cy - dy - gy - hy);
      dzdz = 0.25 * (az + bz + ez + fz - cz - dz - gz - hz);
      aj[p] = 1. / det3(dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz);
    }
  });
  // boundary copies (metrics.c:215-322), x planes, then y, then z
  box(s, 0, 1, 0, my, 0, mz, [=] __device__(int, int j, int k) {
    long p0 = g.ix(0, j, k), p1 = g.ix(1, j, k), q0 = g.ix(mx - 1, j, k), q1 = g.ix(mx - 2, j, k);
    eta.x[p0] = eta.x[p1]; eta.y[p0] = eta.y[p1]; eta.z[p0] = eta.z[p1];
    zet.x[p0] = zet.x[p1]; zet.y[p0] = zet.y[p1]; zet.z[p0] = zet.z[p1]; aj[p0] = aj[p1];
    eta.x[q0] = eta.x[q1]; eta.y[q0] = eta.y[q1]; eta.z[q0] = eta.z[q1];
    zet.x[q0] = zet.x[q1]; zet.y[q0] = zet.y[q1]; zet.z[q0] = zet.z[q1]; aj[q0] = aj[q1];
  });
  box(s, 0, mx, 0, 1, 0, mz, [=] __device__(int i, int, int k) {
    long p0 = g.ix(i, 0, k), p1 = g.ix(i, 1, k), q0 = g.ix(i, my - 1, k), q1 = g.ix(i, my - 2, k);
    csi.x[p0] = csi.x[p1]; csi.y[p0] = csi.y[p1]; csi.z[p0] = csi.z[p1];
    zet.x[p0] = zet.x[p1]; zet.y[p0] = zet.y[p1]; zet.z[p0] = zet.z[p1]; aj[p0] = aj[p1];
    csi.x[q0] = csi.x[q1]; csi.y[q0] = csi.y[q1]; csi.z[q0] = csi.z[q1];
    zet.x[q0] = zet.x[q1]; zet.y[q0] = zet.y[q1]; zet.z[q0] = zet.z[q1]; aj[q0] = aj[q1];
  });
  box(s, 0, mx, 0, my, 0, 1, [=] __device__(int i, int j, int) {
    long p0 = g.ix(i, j, 0), p1 = g.ix(i, j, 1), q0 = g.ix(i, j, mz - 1), q1 = g.ix(i, j, mz - 2);
    eta.x[p0] = eta.x[p1]; eta.y[p0] = eta.y[p1]; eta.z[p0] = eta.z[p1];
    csi.x[p0] = csi.x[p1]; csi.y[p0] = csi.y[p1]; csi.z[p0] = csi.z[p1]; aj[p0] = aj[p1];
    eta.x[q0] = eta.x[q1]; eta.y[q0] = eta.y[q1]; eta.z[q0] = eta.z[q1];
    csi.x[q0] = csi.x[q1]; csi.y[q0] = csi.y[q1]; csi.z[q0] = csi.z[q1]; aj[q0] = aj[q1];
  });

  // ---- grid spacing (metrics.c:351-413) ----
  PV gs = M.gs;
  box(s, 1, mx - 1, 1, my - 1, 1, mz - 1, [=] __device__(int i, int j, int k) {
    long p = g.ix(i, j, k);
    double ax = X(k, j, i), bx = X(k, j - 1, i), cx = X(k - 1, j, i), dx = X(k - 1, j - 1, i);
    double ex = X(k, j, i - 1), fx = X(k, j - 1, i - 1), gx = X(k - 1, j, i - 1), hx = X(k - 1, j - 1, i - 1);
    double ay = Y(k, j, i), by = Y(k, j - 1, i), cy = Y(k - 1, j, i), dy = Y(k - 1, j - 1, i);
    double ey = Y(k, j, i - 1), fy = Y(k, j - 1, i - 1), gy = Y(k - 1, j, i - 1), hy = Y(k - 1, j - 1, i - 1);
    double az = Z(k, j, i), bz = Z(k, j - 1, i), cz = Z(k - 1, j, i), dz = Z(k - 1, j - 1, i);
    double ez = Z(k, j, i - 1), fz = Z(k, j - 1, i - 1), gz = Z(k - 1, j, i - 1), hz = Z(k - 1, j - 1, i - 1);
    double xcp, ycp, zcp, xcm, ycm, zcm;
    xcp = 0.25 * (ax + bx + dx + cx); ycp = 0.25 * (ay + by + dy + cy); zcp = 0.25 * (az + bz + dz + cz);
    xcm = 0.25 * (ex + fx + hx + gx); ycm = 0.25 * (ey + fy + hy + gy); zcm = 0.25 * (ez + fz + hz + gz);
    gs.x[p] = sqrt((xcp - xcm) * (xcp - xcm) + (ycp - ycm) * (ycp - ycm) + (zcp - zcm) * (zcp - zcm));
    xcp = 0.25 * (ax + ex + cx + gx); ycp = 0.25 * (ay + ey + cy + gy); zcp = 0.25 * (az + ez + cz + gz);
    xcm = 0.25 * (bx + fx + dx + hx); ycm = 0.25 * (by + fy + dy + hy); zcm = 0.25 * (bz + fz + dz + hz);
    gs.y[p] = sqrt((xcp - xcm) * (xcp - xcm) + (ycp - ycm) * (ycp - ycm) + (zcp - zcm) * (zcp - zcm));
    xcp = 0.25 * (ax + ex + bx + fx); ycp = 0.25 * (ay + ey + by + fy); zcp = 0.25 * (az + ez + bz + fz);
    xcm = 0.25 * (cx + gx + dx + hx); ycm = 0.25 * (cy + gy + dy + hy); zcm = 0.25 * (cz + gz + dz + hz);
    gs.z[p] = sqrt((xcp - xcm) * (xcp - xcm) + (ycp - ycm) * (ycp - ycm) + (zcp - zcm) * (zcp - zcm));
  });
  g2l(s, g, gs.x); g2l(s, g, gs.y); g2l(s, g, gs.z);

  const int sw = (xper || zper) ? 3 : 2;
  const int gxs = xper ? -sw : 0, gxe = xper ? mx + sw : mx;
  const int gys = 0, gye = my;
  const int gzs = zper ? -sw : 0, gze = zper ? mz + sw : mz;

  // ---- I faces (metrics.c:644-790) ----
  PV cx = M.cx;
  box(s, gxs, gxe, gys + 1, gye, gzs + 1, gze, [=] __device__(int i, int j, int k) {
    if (!(i > -1 && j > 0 && k > 0 && i < IM + 1 && j < JM + 1 && k < KM + 1)) return;
    long p = g.ix(i, j, k);
    cx.x[p] = (X(k, j, i) + X(k - 1, j, i) + X(k, j - 1, i) + X(k - 1, j - 1, i)) * 0.25;
    cx.y[p] = (Y(k, j, i) + Y(k - 1, j, i) + Y(k, j - 1, i) + Y(k - 1, j - 1, i)) * 0.25;
    cx.z[p] = (Z(k, j, i) + Z(k - 1, j, i) + Z(k, j - 1, i) + Z(k - 1, j - 1, i)) * 0.25;
  });
  if (bct0 == 7) {
    box(s, 0, 1, gys + 1, gye, gzs + 1, gze, [=] __device__(int, int j, int k) {
      long pm = g.ix(-1, j, k), p0 = g.ix(0, j, k), q = g.ix(mx - 1, j, k), q1 = g.ix(mx - 2, j, k);
      cx.x[pm] = cx.x[p0] - gs.x[g.ix(-2, j, k)]; cx.y[pm] = cx.y[p0]; cx.z[pm] = cx.z[p0];
      cx.x[q] = cx.x[q1] + gs.x[g.ix(mx + 1, j, k)]; cx.y[q] = cx.y[q1]; cx.z[q] = cx.z[q1];
    });
  }
  PV icsi = M.icsi, ieta = M.ieta, izet = M.izet;
  double *iaj = M.iaj;
  box(s, 0, mx - 1, 1, my - 1, 1, mz - 1, [=] __device__(int i, int j, int k) {
#define CX(a, b, cc) cx.x[g.ix(cc, b, a)]
#define CY(a, b, cc) cx.y[g.ix(cc, b, a)]
#define CZ(a, b, cc) cx.z[g.ix(cc, b, a)]
    double dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz;
    if (i == 0 && bct0 != 7) {
      dxdc = CX(k, j, i + 1) - CX(k, j, i); dydc = CY(k, j, i + 1) - CY(k, j, i); dzdc = CZ(k, j, i + 1) - CZ(k, j, i);
    } else if (i == mx - 2 && bct0 != 7) {
      dxdc = CX(k, j, i) - CX(k, j, i - 1); dydc = CY(k, j, i) - CY(k, j, i - 1); dzdc = CZ(k, j, i) - CZ(k, j, i - 1);
    } else {
      dxdc = (CX(k, j, i + 1) - CX(k, j, i - 1)) * 0.5; dydc = (CY(k, j, i + 1) - CY(k, j, i - 1)) * 0.5;
      dzdc = (CZ(k, j, i + 1) - CZ(k, j, i - 1)) * 0.5;
    }
    if (j > 1 && j < my - 2) {
      dxde = (CX(k, j + 1, i) - CX(k, j - 1, i)) * 0.5; dyde = (CY(k, j + 1, i) - CY(k, j - 1, i)) * 0.5;
      dzde = (CZ(k, j + 1, i) - CZ(k, j - 1, i)) * 0.5;
    } else if (j == 1) {
      dxde = CX(k, j + 1, i) - CX(k, j, i); dyde = CY(k, j + 1, i) - CY(k, j, i); dzde = CZ(k, j + 1, i) - CZ(k, j, i);
    } else {
      dxde = CX(k, j, i) - CX(k, j - 1, i); dyde = CY(k, j, i) - CY(k, j - 1, i); dzde = CZ(k, j, i) - CZ(k, j - 1, i);
    }
    if (k > 1 && k < mz - 2) {
      dxdz = (CX(k + 1, j, i) - CX(k - 1, j, i)) * 0.5; dydz = (CY(k + 1, j, i) - CY(k - 1, j, i)) * 0.5;
      dzdz = (CZ(k + 1, j, i) - CZ(k - 1, j, i)) * 0.5;
    } else if (k == 1) {
      dxdz = CX(k + 1, j, i) - CX(k, j, i); dydz = CY(k + 1, j, i) - CY(k, j, i); dzdz = CZ(k + 1, j, i) - CZ(k, j, i);
    } else {
      dxdz = CX(k, j, i) - CX(k - 1, j, i); dydz = CY(k, j, i) - CY(k - 1, j, i); dzdz = CZ(k, j, i) - CZ(k - 1, j, i);
    }
#undef CX
#undef CY
#undef CZ
    long p = g.ix(i, j, k);
    cross3(dxde, dyde, dzde, dxdz, dydz, dzdz, &icsi.x[p], &icsi.y[p], &icsi.z[p]);
    cross3(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &ieta.x[p], &ieta.y[p], &ieta.z[p]);
    cross3(dxdc, dydc, dzdc, dxde, dyde, dzde, &izet.x[p], &izet.y[p], &izet.z[p]);
    iaj[p] = 1. / det3(dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz);
  });

  // ---- J faces (metrics.c:806-942) ----
  PV cy = M.cy;
  box(s, gxs + 1, gxe, gys, gye, gzs + 1, gze, [=] __device__(int i, int j, int k) {
    if (!(i > 0 && j > -1 && k > 0 && i < IM + 1 && j < JM + 1 && k < KM + 1)) return;
    long p = g.ix(i, j, k);
    cy.x[p] = (X(k, j, i) + X(k - 1, j, i) + X(k, j, i - 1) + X(k - 1, j, i - 1)) * 0.25;
    cy.y[p] = (Y(k, j, i) + Y(k - 1, j, i) + Y(k, j, i - 1) + Y(k - 1, j, i - 1)) * 0.25;
    cy.z[p] = (Z(k, j, i) + Z(k - 1, j, i) + Z(k, j, i - 1) + Z(k - 1, j, i - 1)) * 0.25;
  });
  PV jcsi = M.jcsi, jeta = M.jeta, jzet = M.jzet;
  double *jaj = M.jaj;
  box(s, 1, mx - 1, 0, my - 1, 1, mz - 1, [=] __device__(int i, int j, int k) {
#define CX(a, b, cc) cy.x[g.ix(cc, b, a)]
#define CY(a, b, cc) cy.y[g.ix(cc, b, a)]
#define CZ(a, b, cc) cy.z[g.ix(cc, b, a)]
    double dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz;
    if (i > 1 && i < mx - 2) {
      dxdc = (CX(k, j, i + 1) - CX(k, j, i - 1)) * 0.5; dydc = (CY(k, j, i + 1) - CY(k, j, i - 1)) * 0.5;
      dzdc = (CZ(k, j, i + 1) - CZ(k, j, i - 1)) * 0.5;
    } else if (i == 1) {
      dxdc = CX(k, j, i + 1) - CX(k, j, i); dydc = CY(k, j, i + 1) - CY(k, j, i); dzdc = CZ(k, j, i + 1) - CZ(k, j, i);
    } else {
      dxdc = CX(k, j, i) - CX(k, j, i - 1); dydc = CY(k, j, i) - CY(k, j, i - 1); dzdc = CZ(k, j, i) - CZ(k, j, i - 1);
    }
    if (j == 0) {
      dxde = CX(k, j + 1, i) - CX(k, j, i); dyde = CY(k, j + 1, i) - CY(k, j, i); dzde = CZ(k, j + 1, i) - CZ(k, j, i);
    } else if (j == my - 2) {
      dxde = CX(k, j, i) - CX(k, j - 1, i); dyde = CY(k, j, i) - CY(k, j - 1, i); dzde = CZ(k, j, i) - CZ(k, j - 1, i);
    } else {
      dxde = (CX(k, j + 1, i) - CX(k, j - 1, i)) * 0.5; dyde = (CY(k, j + 1, i) - CY(k, j - 1, i)) * 0.5;
      dzde = (CZ(k, j + 1, i) - CZ(k, j - 1, i)) * 0.5;
    }
    if (k > 1 && k < mz - 2) {
      dxdz = (CX(k + 1, j, i) - CX(k - 1, j, i)) * 0.5; dydz = (CY(k + 1, j, i) - CY(k - 1, j, i)) * 0.5;
      dzdz = (CZ(k + 1, j, i) - CZ(k - 1, j, i)) * 0.5;
    } else if (k == 1) {
      dxdz = CX(k + 1, j, i) - CX(k, j, i); dydz = CY(k + 1, j, i) - CY(k, j, i); dzdz = CZ(k + 1, j, i) - CZ(k, j, i);
    } else {
      dxdz = CX(k, j, i) - CX(k - 1, j, i); dydz = CY(k, j, i) - CY(k - 1, j, i); dzdz = CZ(k, j, i) - CZ(k - 1, j, i);
    }
#undef CX
#undef CY
#undef CZ
    long p = g.ix(i, j, k);
    cross3(dxde, dyde, dzde, dxdz, dydz, dzdz, &jcsi.x[p], &jcsi.y[p], &jcsi.z[p]);
    cross3(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &jeta.x[p], &jeta.y[p], &jeta.z[p]);
    cross3(dxdc, dydc, dzdc, dxde, dyde, dzde, &jzet.x[p], &jzet.y[p], &jzet.z[p]);
    jaj[p] = 1. / det3(dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz);
  });

  // ---- K faces (metrics.c:956-1090) ----
  PV cz = M.cz;
  box(s, gxs + 1, gxe, gys + 1, gye, gzs, gze, [=] __device__(int i, int j, int k) {
    if (!(i > 0 && j > 0 && k > -1 && i < IM + 1 && j < JM + 1 && k < KM + 1)) return;
    long p = g.ix(i, j, k);
    cz.x[p] = (X(k, j, i) + X(k, j - 1, i) + X(k, j, i - 1) + X(k, j - 1, i - 1)) * 0.25;
    cz.y[p] = (Y(k, j, i) + Y(k, j - 1, i) + Y(k, j, i - 1) + Y(k, j - 1, i - 1)) * 0.25;
    cz.z[p] = (Z(k, j, i) + Z(k, j - 1, i) + Z(k, j, i - 1) + Z(k, j - 1, i - 1)) * 0.25;
  });
  if (bct4 == 7) {
    box(s, gxs + 1, gxe, gys + 1, gye, 0, 1, [=] __device__(int i, int j, int) {
      long pm = g.ix(i, j, -1), p0 = g.ix(i, j, 0), q = g.ix(i, j, mz - 1), q1 = g.ix(i, j, mz - 2);
      cz.x[pm] = cz.x[p0]; cz.y[pm] = cz.y[p0]; cz.z[pm] = cz.z[p0] - gs.z[g.ix(i, j, -2)];
      cz.x[q] = cz.x[q1]; cz.y[q] = cz.y[q1]; cz.z[q] = cz.z[q1] + gs.z[g.ix(i, j, mz + 1)];
    });
  }
  PV kcsi = M.kcsi, keta = M.keta, kzet = M.kzet;
  double *kaj = M.kaj;
  box(s, 1, mx - 1, 1, my - 1, 0, mz - 1, [=] __device__(int i, int j, int k) {
#define CX(a, b, cc) cz.x[g.ix(cc, b, a)]
#define CY(a, b, cc) cz.y[g.ix(cc, b, a)]
#define CZ(a, b, cc) cz.z[g.ix(cc, b, a)]
    double dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz;
    if (i > 1 && i < mx - 2) {
      dxdc = (CX(k, j, i + 1) - CX(k, j, i - 1)) * 0.5; dydc = (CY(k, j, i + 1) - CY(k, j, i - 1)) * 0.5;
      dzdc = (CZ(k, j, i + 1) - CZ(k, j, i - 1)) * 0.5;
    } else if (i == 1) {
      dxdc = CX(k, j, i + 1) - CX(k, j, i); dydc = CY(k, j, i + 1) - CY(k, j, i); dzdc = CZ(k, j, i + 1) - CZ(k, j, i);
    } else {
      dxdc = CX(k, j, i) - CX(k, j, i - 1); dydc = CY(k, j, i) - CY(k, j, i - 1); dzdc = CZ(k, j, i) - CZ(k, j, i - 1);
    }
    if (j > 1 && j < my - 2) {
      dxde = (CX(k, j + 1, i) - CX(k, j - 1, i)) * 0.5; dyde = (CY(k, j + 1, i) - CY(k, j - 1, i)) * 0.5;
      dzde = (CZ(k, j + 1, i) - CZ(k, j - 1, i)) * 0.5;
    } else if (j == 1) {
      dxde = CX(k, j + 1, i) - CX(k, j, i); dyde = CY(k, j + 1, i) - CY(k, j, i); dzde = CZ(k, j + 1, i) - CZ(k, j, i);
    } else {
      dxde = CX(k, j, i) - CX(k, j - 1, i); dyde = CY(k, j, i) - CY(k, j - 1, i); dzde = CZ(k, j, i) - CZ(k, j - 1, i);
    }
    if (k == 0 && bct4 != 7) {
      dxdz = CX(k + 1, j, i) - CX(k, j, i); dydz = CY(k + 1, j, i) - CY(k, j, i); dzdz = CZ(k + 1, j, i) - CZ(k, j, i);
    } else if (k == mz - 2 && bct4 != 7) {
      dxdz = CX(k, j, i) - CX(k - 1, j, i); dydz = CY(k, j, i) - CY(k - 1, j, i); dzdz = CZ(k, j, i) - CZ(k - 1, j, i);
    } else {
      dxdz = (CX(k + 1, j, i) - CX(k - 1, j, i)) * 0.5; dydz = (CY(k + 1, j, i) - CY(k - 1, j, i)) * 0.5;
      dzdz = (CZ(k + 1, j, i) - CZ(k - 1, j, i)) * 0.5;
    }
#undef CX
#undef CY
#undef CZ
    long p = g.ix(i, j, k);
    cross3(dxde, dyde, dzde, dxdz, dydz, dzdz, &kcsi.x[p], &kcsi.y[p], &kcsi.z[p]);
    cross3(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &keta.x[p], &keta.y[p], &keta.z[p]);
    cross3(dxdc, dydc, dzdc, dxde, dyde, dzde, &kzet.x[p], &kzet.y[p], &kzet.z[p]);
    kaj[p] = 1. / det3(dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz);
  });
#undef X
#undef Y
#undef Z

  // ---- global -> local, then Update_Metrics_PBC ----
  double *all[] = {csi.x, csi.y, csi.z, eta.x, eta.y, eta.z, zet.x, zet.y, zet.z, aj,
                   icsi.x, icsi.y, icsi.z, ieta.x, ieta.y, ieta.z, izet.x, izet.y, izet.z, iaj,
                   jcsi.x, jcsi.y, jcsi.z, jeta.x, jeta.y, jeta.z, jzet.x, jzet.y, jzet.z, jaj,
                   kcsi.x, kcsi.y, kcsi.z, keta.x, keta.y, keta.z, kzet.x, kzet.y, kzet.z, kaj};
  const int nall = sizeof(all) / sizeof(all[0]);
  for (int q = 0; q < nall; q++) g2l(s, g, all[q]);
  for (int q = 0; q < nall; q++) {
    double *a = all[q];
    if (bct0 == 7)
      box(s, 0, 1, gys, gye, gzs, gze, [=] __device__(int, int j, int k) {
        a[g.ix(0, j, k)] = a[g.ix(-2, j, k)];
        a[g.ix(-1, j, k)] = a[g.ix(-3, j, k)];
        a[g.ix(mx - 1, j, k)] = a[g.ix(mx + 1, j, k)];
        a[g.ix(mx, j, k)] = a[g.ix(mx + 2, j, k)];
      });
  }
  for (int q = 0; q < nall; q++) {
    double *a = all[q];
    if (bct4 == 7)
      box(s, gxs, gxe, gys, gye, 0, 1, [=] __device__(int i, int j, int) {
        a[g.ix(i, j, 0)] = a[g.ix(i, j, -2)];
        a[g.ix(i, j, -1)] = a[g.ix(i, j, -3)];
        a[g.ix(i, j, mz - 1)] = a[g.ix(i, j, mz + 1)];
        a[g.ix(i, j, mz)] = a[g.ix(i, j, mz + 2)];
      });
  }

  // ---- compact copies: operator coefficients for every level, the full
  // metric set for the finest ----
  {
    double *gg[9];
    for (int q = 0; q < 9; q++) gg[q] = L.g[q];
    double *ajc = L.aj;
    double *g0 = gg[0], *g1 = gg[1], *g2 = gg[2], *g3 = gg[3], *g4 = gg[4], *g5 = gg[5], *g6 = gg[6],
           *g7 = gg[7], *g8 = gg[8];
    box(s, 0, mx, 0, my, 0, mz, [=] __device__(int i, int j, int k) {
      long p = g.ix(i, j, k), q = (long)i + (long)mx * (j + (long)my * k);
      // g coefficients (poisson.c:699-735)
      double cx_ = icsi.x[p], cy_ = icsi.y[p], cz_ = icsi.z[p];
      double ex_ = ieta.x[p], ey_ = ieta.y[p], ez_ = ieta.z[p];
      double zx_ = izet.x[p], zy_ = izet.y[p], zz_ = izet.z[p];
      double a = iaj[p];
      g0[q] = (cx_ * cx_ + cy_ * cy_ + cz_ * cz_) * a;
      g1[q] = (ex_ * cx_ + ey_ * cy_ + ez_ * cz_) * a;
      g2[q] = (zx_ * cx_ + zy_ * cy_ + zz_ * cz_) * a;
      cx_ = jcsi.x[p]; cy_ = jcsi.y[p]; cz_ = jcsi.z[p];
      ex_ = jeta.x[p]; ey_ = jeta.y[p]; ez_ = jeta.z[p];
      zx_ = jzet.x[p]; zy_ = jzet.y[p]; zz_ = jzet.z[p];
      a = jaj[p];
      g3[q] = (cx_ * ex_ + cy_ * ey_ + cz_ * ez_) * a;
      g4[q] = (ex_ * ex_ + ey_ * ey_ + ez_ * ez_) * a;
      g5[q] = (zx_ * ex_ + zy_ * ey_ + zz_ * ez_) * a;
      cx_ = kcsi.x[p]; cy_ = kcsi.y[p]; cz_ = kcsi.z[p];
      ex_ = keta.x[p]; ey_ = keta.y[p]; ez_ = keta.z[p];
      zx_ = kzet.x[p]; zy_ = kzet.y[p]; zz_ = kzet.z[p];
      a = kaj[p];
      g6[q] = (cx_ * zx_ + cy_ * zy_ + cz_ * zz_) * a;
      g7[q] = (ex_ * zx_ + ey_ * zy_ + ez_ * zz_) * a;
      g8[q] = (zx_ * zx_ + zy_ * zy_ + zz_ * zz_) * a;
      ajc[q] = aj[p];
    });
    if (l == c->nlev - 1) {
      CvMetrics &Mc = c->m;
      double *dst[] = {Mc.csi[0], Mc.csi[1], Mc.csi[2], Mc.eta[0], Mc.eta[1], Mc.eta[2],
                       Mc.zet[0], Mc.zet[1], Mc.zet[2], Mc.aj,
                       Mc.icsi[0], Mc.icsi[1], Mc.icsi[2], Mc.ieta[0], Mc.ieta[1], Mc.ieta[2],
                       Mc.izet[0], Mc.izet[1], Mc.izet[2], Mc.iaj,
                       Mc.jcsi[0], Mc.jcsi[1], Mc.jcsi[2], Mc.jeta[0], Mc.jeta[1], Mc.jeta[2],
                       Mc.jzet[0], Mc.jzet[1], Mc.jzet[2], Mc.jaj,
                       Mc.kcsi[0], Mc.kcsi[1], Mc.kcsi[2], Mc.keta[0], Mc.keta[1], Mc.keta[2],
                       Mc.kzet[0], Mc.kzet[1], Mc.kzet[2], Mc.kaj};
      for (int q = 0; q < nall; q++) {
        double *d = dst[q];
        const double *a = all[q];
        box(s, 0, mx, 0, my, 0, mz, [=] __device__(int i, int j, int k) {
          d[(long)i + (long)mx * (j + (long)my * k)] = a[g.ix(i, j, k)];
        });
      }
    }
  }
  CV_CHECK(hipStreamSynchronize(s));
  CV_CHECK(hipFree(buf));
  sbuf.p = nullptr;
  return CV_OK;
}

// the device scratch cv_set_geometry holds while it builds the finest level
// (the largest): the padded temporaries plus the node coordinates
long cv_geometry_scratch_bytes(const cv_grid_desc *d) {
  const long px = d->IM + 1 + 2 * PG, py = d->JM + 1 + 2 * PG, pz = d->KM + 1 + 2 * PG;
  return (long)(3 * CV_METRIC_SCRATCH_VEC + CV_METRIC_SCRATCH_SC) * px * py * pz * (long)sizeof(double) +
         3L * d->IM * d->JM * d->KM * (long)sizeof(double);
}

int cv_metrics_build(cv_ctx *c, const double *xyz_host) {
  for (int l = c->nlev - 1; l >= 0; l--) {
    int rc = level_metrics(c, l, xyz_host);
    if (rc) return rc;
  }
  return CV_OK;
}

extern "C" int cv_set_geometry(cv_ctx *c, const double *xyz) {
  if (!c || !xyz) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(c->device));
  int rc = cv_metrics_build(c, xyz);
  if (rc) return rc;
  c->geometry_set = 1;
  if (c->nvert_set) return cv_poisson_prepare(c);
  return CV_OK;
}
