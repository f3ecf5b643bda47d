// cv_momentum.hip -- momentum right-hand side and the fractional-step
// driver on gfx950.
//
//  Contra2Cart        rhs.c:25-144
//  InflowFlux         bcs.c:99-246     (bctype 5 on the k- face, analytic profiles)
//  OutflowFlux        bcs.c:403-469    (bctype 4 on the k+ face)
//  FormBCS            bcs.c:4436-5888  (bctype 1 wall, 3 slip, 6 far field, 7 periodic,
//                                       5 inlet, 4 outlet)
//  GhostNodeVelocity  bcs.c:3645-4435  (same subset)
//  FormFunction1      rhs.c:2567-3351 with Convection (QUICK, rhs.c:1463-1945)
//                     and Viscous (laminar, rhs.c:1947-2565)
//  RungeKutta         implicitsolver.c:32-119
//  Flow_Solver        solvers.c:944-1156
//
// Each face flux (Visc - Conv) is evaluated once per face into a face
// family array, as the reference's Fp1/Fp2/Fp3 temporaries; the arithmetic
// per face is the reference's.
#include "cv_internal.h"
#include <cstring>
#include <algorithm>
#include "cv_reduce.h"
#include <cmath>

#define BS 256

namespace {

struct G {
  int mx, my, mz, xper, zper;
  long sj, sk;
  __device__ long at(int i, int j, int k) const { return (long)i + (long)mx * ((long)j + (long)my * k); }
  // storage index -> (i, j, k); storage sizes stay below 2^31 (cv_ctx_create)
  __device__ void dec(long t, int &i, int &j, int &k) const {
    unsigned u = (unsigned)t;
    i = (int)(u % (unsigned)mx);
    u /= (unsigned)mx;
    j = (int)(u % (unsigned)my);
    k = (int)(u / (unsigned)my);
  }
  // local-array index of the reference's ghosted vectors for the QUICK far
  // ghosts (-1 -> m-3, m -> 2 on a periodic axis, rhs.c:1557-1628)
  __device__ int fi(int i) const { return xper ? (i == -1 ? mx - 3 : (i == mx ? 2 : i)) : i; }
  __device__ int fk(int k) const { return zper ? (k == -1 ? mz - 3 : (k == mz ? 2 : k)) : k; }
  __device__ long atq(int i, int j, int k) const { return at(fi(i), j, fk(k)); }
};

struct V3 { double *x, *y, *z; };
struct C3 { const double *x, *y, *z; };

inline unsigned nblk(long n) {
  long b = (n + BS - 1) / BS;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// k-marching launch shape for the cell and face kernels: a block of
// 64 x MARCH_TY threads owns an (i, j) tile and walks MARCH_KC planes of k,
// so the k-1 .. k+2 neighbours a cell reads were loaded by the same block a
// few iterations earlier and are still in the cache
#define MARCH_TY 4
#define MARCH_KC 16
#define CV_MARCH(g)                                                                          \
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * MARCH_TY + threadIdx.y;      \
  const int k0_ = blockIdx.z * MARCH_KC, k1_ = min(k0_ + MARCH_KC, (g).mz);                 \
  if (i < (g).mx && j < (g).my)                                                              \
    for (int k = k0_; k < k1_; k++)

// the same launch shape with the block given explicitly (batched launches)
#define CV_MARCH_B(g, bx_, by_, bz_)                                                          \
  const int i = (bx_) * 64 + threadIdx.x, j = (by_) * MARCH_TY + threadIdx.y;               \
  const int k0_ = (bz_) * MARCH_KC, k1_ = min(k0_ + MARCH_KC, (g).mz);                      \
  if (i < (g).mx && j < (g).my)                                                              \
    for (int k = k0_; k < k1_; k++)

// a grid-stride partition: first index and stride of this thread
struct GS {
  long t0, step;
};
#define GS_SELF GS{blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x}

inline dim3 march_grid(int mx, int my, int mz) {
  return dim3((unsigned)((mx + 63) / 64), (unsigned)((my + MARCH_TY - 1) / MARCH_TY),
              (unsigned)((mz + MARCH_KC - 1) / MARCH_KC));
}
#define MARCH_BLOCK dim3(64, MARCH_TY)

// ------------------------------------------------------------------
// Contra2Cart (rhs.c:25-144)
// ------------------------------------------------------------------
// Cramer's rule for Ucat of one cell from its averaged face vectors (rows
// m0, m1, m2) and averaged fluxes q (rhs.c:25-144)
__device__ __forceinline__ void c2c_solve(double m00, double m01, double m02, double m10, double m11, double m12,
                                          double m20, double m21, double m22, double q0, double q1, double q2, V3 ucat,
                                          long t) {
  double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
  double det0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
  double det1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
  double det2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
  ucat.x[t] = det0 / det;
  ucat.y[t] = det1 / det;
  ucat.z[t] = det2 / det;
}

// one interior cell (i, j, k); IB cells (0.1 <= nvert <= 7) keep their value
__device__ __forceinline__ void c2c_cell(G g, C3 uc, C3 csi, C3 eta, C3 zet, const double *nv, V3 ucat, int i, int j, int k) {
  const long t = g.at(i, j, k);
  const double nn = nv[t];
  if (!(nn < 0.1 || nn > 7.)) return;
  const long ti = t - 1, tj = t - g.sj, tk = t - g.sk;
  c2c_solve(0.5 * (csi.x[ti] + csi.x[t]), 0.5 * (csi.y[ti] + csi.y[t]), 0.5 * (csi.z[ti] + csi.z[t]),
            0.5 * (eta.x[tj] + eta.x[t]), 0.5 * (eta.y[tj] + eta.y[t]), 0.5 * (eta.z[tj] + eta.z[t]),
            0.5 * (zet.x[tk] + zet.x[t]), 0.5 * (zet.y[tk] + zet.y[t]), 0.5 * (zet.z[tk] + zet.z[t]),
            0.5 * (uc.x[ti] + uc.x[t]), 0.5 * (uc.y[tj] + uc.y[t]), 0.5 * (uc.z[tk] + uc.z[t]), ucat, t);
}
// the k-march of the interior: the k-1 plane's zet and Ucont.z a cell reads
// are the previous iteration's own values, carried in registers (a whole
// plane of the XCD's blocks outgrows its L2 between the two reads)
__device__ __forceinline__ void contra2cart_body(G g, C3 uc, C3 csi, C3 eta, C3 zet, const double *nv, V3 ucat, long n, int bx, int by, int bz) {
  (void)n;
  const int i = bx * 64 + threadIdx.x, j = by * MARCH_TY + threadIdx.y;
  if (!(i > 0 && i < g.mx - 1 && j > 0 && j < g.my - 1)) return;
  const int k0 = max(bz * MARCH_KC, 1), k1 = min(bz * MARCH_KC + MARCH_KC, g.mz - 1);
  if (k0 >= k1) return;
  long t = g.at(i, j, k0 - 1);
  double zx = zet.x[t], zy = zet.y[t], zz = zet.z[t], wz = uc.z[t];
  for (int k = k0; k < k1; k++) {
    t += g.sk;
    const double zx1 = zet.x[t], zy1 = zet.y[t], zz1 = zet.z[t], wz1 = uc.z[t];
    const double nn = nv[t];
    if (nn < 0.1 || nn > 7.) {
      const long ti = t - 1, tj = t - g.sj;
      c2c_solve(0.5 * (csi.x[ti] + csi.x[t]), 0.5 * (csi.y[ti] + csi.y[t]), 0.5 * (csi.z[ti] + csi.z[t]),
                0.5 * (eta.x[tj] + eta.x[t]), 0.5 * (eta.y[tj] + eta.y[t]), 0.5 * (eta.z[tj] + eta.z[t]),
                0.5 * (zx + zx1), 0.5 * (zy + zy1), 0.5 * (zz + zz1),
                0.5 * (uc.x[ti] + uc.x[t]), 0.5 * (uc.y[tj] + uc.y[t]), 0.5 * (wz + wz1), ucat, t);
    }
    zx = zx1; zy = zy1; zz = zz1; wz = wz1;
  }
}

// One k-march for an RK stage on faces without an inlet, far-field, slip
// or outlet (FormBCS then changes only ghosts): U = U_o + a Rhs on every
// point, the wall-normal flux reset of wall_flux_zero_body (masks wx, wy,
// wz), and Contra2Cart of the updated U on the interior.  The updated
// fluxes of the i-1 and j-1 neighbours are formed again from U_o and Rhs
// with the same expression, the k-1 plane's is carried, so the bits are
// those of the three passes in turn.
__device__ __forceinline__ void rk_c2c_body(G g, double a, C3 uo, C3 r, V3 u, int wx, int wy, int wz, C3 csi, C3 eta,
                                            C3 zet, const double *nv, V3 ucat, int bx, int by, int bz) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  const int i = bx * 64 + threadIdx.x, j = by * MARCH_TY + threadIdx.y;
  if (!(i < mx && j < my)) return;
  const int k0 = bz * MARCH_KC, k1 = min(k0 + MARCH_KC, mz);
  auto ux = [&](long q, int ii) -> double {
    const double v = a * r.x[q] + uo.x[q];
    return (((wx & 1) && ii == 0) || ((wx & 2) && ii == mx - 2)) ? 0. : v;
  };
  auto uy = [&](long q, int jj) -> double {
    const double v = a * r.y[q] + uo.y[q];
    return (((wy & 1) && jj == 0) || ((wy & 2) && jj == my - 2)) ? 0. : v;
  };
  auto uz = [&](long q, int kk) -> double {
    const double v = a * r.z[q] + uo.z[q];
    return (((wz & 1) && kk == 0) || ((wz & 2) && kk == mz - 2)) ? 0. : v;
  };
  const bool in_ij = i > 0 && i < mx - 1 && j > 0 && j < my - 1;
  long t = g.at(i, j, k0);
  // the k-1 plane's zet and updated Ucont.z (carried from one plane to the next)
  double zx = 0., zy = 0., zz = 0., pz = 0.;
  if (in_ij && k0 >= 1) {
    const long q = t - g.sk;
    zx = zet.x[q]; zy = zet.y[q]; zz = zet.z[q];
    pz = uz(q, k0 - 1);
  }
  for (int k = k0; k < k1; k++, t += g.sk) {
    const double vx = ux(t, i), vy = uy(t, j), vz = uz(t, k);
    u.x[t] = vx; u.y[t] = vy; u.z[t] = vz;
    if (!in_ij) continue;
    const double zx1 = zet.x[t], zy1 = zet.y[t], zz1 = zet.z[t];
    if (k > 0 && k < mz - 1) {
      const double nn = nv[t];
      if (nn < 0.1 || nn > 7.) {
        const long ti = t - 1, tj = t - g.sj;
        c2c_solve(0.5 * (csi.x[ti] + csi.x[t]), 0.5 * (csi.y[ti] + csi.y[t]), 0.5 * (csi.z[ti] + csi.z[t]),
                  0.5 * (eta.x[tj] + eta.x[t]), 0.5 * (eta.y[tj] + eta.y[t]), 0.5 * (eta.z[tj] + eta.z[t]),
                  0.5 * (zx + zx1), 0.5 * (zy + zy1), 0.5 * (zz + zz1),
                  0.5 * (ux(ti, i - 1) + vx), 0.5 * (uy(tj, j - 1) + vy), 0.5 * (pz + vz), ucat, t);
      }
    }
    zx = zx1; zy = zy1; zz = zz1; pz = vz;
  }
}
__global__ void k_contra2cart(G g, C3 uc, C3 csi, C3 eta, C3 zet, const double *nv, V3 ucat, long n) { contra2cart_body(g, uc, csi, eta, zet, nv, ucat, n, blockIdx.x, blockIdx.y, blockIdx.z); }
__global__ void __launch_bounds__(64 * MARCH_TY) k_rk_c2c(G g, double a, C3 uo, C3 r, V3 u, int wx, int wy, int wz, C3 csi,
                                                         C3 eta, C3 zet, const double *nv, V3 ucat) {
  rk_c2c_body(g, a, uo, r, u, wx, wy, wz, csi, eta, zet, nv, ucat, blockIdx.x, blockIdx.y, blockIdx.z);
}
// Contra2Cart of the interior cells of one k plane
__global__ void k_contra2cart_plane(G g, C3 uc, C3 csi, C3 eta, C3 zet, const double *nv, V3 ucat, int k) {
  const long m2 = (long)(g.mx - 2) * (g.my - 2);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < m2; t += (long)gridDim.x * blockDim.x)
    c2c_cell(g, uc, csi, eta, zet, nv, ucat, 1 + (int)(t % (g.mx - 2)), 1 + (int)(t / (g.mx - 2)), k);
}


// ------------------------------------------------------------------
// ghost planes (FormBCS bcs.c:5252-5731, GhostNodeVelocity bcs.c:4015-4415)
// ------------------------------------------------------------------
// Ubcs of the six boundary planes in one array: face f (0 i-, 1 i+, 2 j-,
// 3 j+, 4 k-, 5 k+), component q (x, y, z), at the plane index r of the
// face's ghost point (i faces j + my*k, j faces i + mx*k, k faces i + mx*j)
__device__ __forceinline__ long ubf_at(const G &g, int f, int q, long r) {
  const long nyz = (long)g.my * g.mz, nxz = (long)g.mx * g.mz, nxy = (long)g.mx * g.my;
  const long sz = f < 2 ? nyz : (f < 4 ? nxz : nxy);
  const long base = f < 2 ? 3 * f * nyz : (f < 4 ? 6 * nyz + 3 * (f - 2) * nxz : 6 * (nyz + nxz) + 3 * (f - 4) * nxy);
  return base + q * sz + r;
}

// wall ghosts: ucat_g = 2*ubcs - ucat_inner, with Ubcs from ubf (0 on
// walls) or 0 when ubf is null.  full: whole tangential ranges
// (GhostNodeVelocity); otherwise interior ranges (FormBCS)
__device__ __forceinline__ void wall_ghost_body(G g, int bx, int bz, int full, V3 u, const double *ubf, GS gs) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  const int a0 = full ? 0 : 1;
  const long nyz = (long)my * mz, nxz = (long)mx * mz, nxy = (long)mx * my;
  const long tot = nyz + nxz + nxy;
  for (long t = gs.t0; t < tot; t += gs.step) {
    long lo, hi, ln, in_lo, in_hi;
    double lx = 0.0, ly = 0.0, lz = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;   // Ubcs of the two ghosts
    long r;
    int f0;
    if (t < nyz) {                       // x planes
      if (bx) continue;
      r = t; f0 = 0;
      int j = (int)(t % my), k = (int)(t / my);
      if (j < a0 || j >= (full ? my : my - 1) || k < a0 || k >= (full ? mz : mz - 1)) continue;
      lo = g.at(0, j, k); in_lo = g.at(1, j, k); hi = g.at(mx - 1, j, k); in_hi = g.at(mx - 2, j, k);
      (void)ln;
    } else if (t < nyz + nxz) {          // y planes
      r = t - nyz; f0 = 2;
      int i = (int)(r % mx), k = (int)(r / mx);
      if (i < a0 || i >= (full ? mx : mx - 1) || k < a0 || k >= (full ? mz : mz - 1)) continue;
      lo = g.at(i, 0, k); in_lo = g.at(i, 1, k); hi = g.at(i, my - 1, k); in_hi = g.at(i, my - 2, k);
    } else {                             // z planes
      if (bz) continue;
      r = t - nyz - nxz; f0 = 4;
      int i = (int)(r % mx), j = (int)(r / mx);
      if (i < a0 || i >= (full ? mx : mx - 1) || j < a0 || j >= (full ? my : my - 1)) continue;
      lo = g.at(i, j, 0); in_lo = g.at(i, j, 1); hi = g.at(i, j, mz - 1); in_hi = g.at(i, j, mz - 2);
    }
    if (ubf) {
      lx = ubf[ubf_at(g, f0, 0, r)]; ly = ubf[ubf_at(g, f0, 1, r)]; lz = ubf[ubf_at(g, f0, 2, r)];
      hx = ubf[ubf_at(g, f0 + 1, 0, r)]; hy = ubf[ubf_at(g, f0 + 1, 1, r)]; hz = ubf[ubf_at(g, f0 + 1, 2, r)];
    }
    u.x[lo] = 2 * lx - u.x[in_lo]; u.y[lo] = 2 * ly - u.y[in_lo]; u.z[lo] = 2 * lz - u.z[in_lo];
    u.x[hi] = 2 * hx - u.x[in_hi]; u.y[hi] = 2 * hy - u.y[in_hi]; u.z[hi] = 2 * hz - u.z[in_hi];
  }
}
__global__ void k_wall_ghost(G g, int bx, int bz, int full, V3 u, const double *ubf) { wall_ghost_body(g, bx, bz, full, u, ubf, GS_SELF); }


// periodic plane copies restricted to the interior of the other axes
// (bcs.c:5325-5431): all four planes read only interior cells
__device__ __forceinline__ void periodic_planes_body(G g, V3 u, double *p, double *nv, GS gs) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  const long nyz = (long)my * mz, nxy = (long)mx * my;
  for (long t = gs.t0; t < nyz + nxy; t += gs.step) {
    long d0, s0, d1, s1;
    if (t < nyz) {
      if (!g.xper) continue;
      int j = (int)(t % my), k = (int)(t / my);
      if (!(j > 0 && k > 0 && j < my - 1 && k < mz - 1)) continue;
      d0 = g.at(0, j, k); s0 = g.at(mx - 2, j, k); d1 = g.at(mx - 1, j, k); s1 = g.at(1, j, k);
    } else {
      if (!g.zper) continue;
      long r = t - nyz;
      int i = (int)(r % mx), j = (int)(r / mx);
      if (!(i > 0 && j > 0 && i < mx - 1 && j < my - 1)) continue;
      d0 = g.at(i, j, 0); s0 = g.at(i, j, mz - 2); d1 = g.at(i, j, mz - 1); s1 = g.at(i, j, 1);
    }
    u.x[d0] = u.x[s0]; u.y[d0] = u.y[s0]; u.z[d0] = u.z[s0];
    u.x[d1] = u.x[s1]; u.y[d1] = u.y[s1]; u.z[d1] = u.z[s1];
    if (p) { p[d0] = p[s0]; p[d1] = p[s1]; nv[d0] = nv[s0]; nv[d1] = nv[s1]; }
  }
}
__global__ void k_periodic_planes(G g, V3 u, double *p, double *nv) { periodic_planes_body(g, u, p, nv, GS_SELF); }


// edge averages, twelve sequential sub-sweeps (bcs.c:5440-5566); one block
__device__ __forceinline__ void edge_average_body(G g, V3 u, double *p) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  for (int sb = 0; sb < 12; sb++) {
    int len;
    int side = (sb < 8) ? (sb / 4) : ((sb - 8) / 2);
    if (sb < 8) len = ((sb % 4) < 2) ? my : mx;
    else len = mz;
    for (int q = threadIdx.x; q < len; q += blockDim.x) {
      long d, a, b;
      if (sb < 8) {
        int k = side ? mz - 1 : 0, kn = side ? mz - 2 : 1;
        switch (sb % 4) {
          case 0: d = g.at(0, q, k); a = g.at(0, q, kn); b = g.at(1, q, k); break;
          case 1: d = g.at(mx - 1, q, k); a = g.at(mx - 1, q, kn); b = g.at(mx - 2, q, k); break;
          case 2: d = g.at(q, 0, k); a = g.at(q, 0, kn); b = g.at(q, 1, k); break;
          default: d = g.at(q, my - 1, k); a = g.at(q, my - 1, kn); b = g.at(q, my - 2, k); break;
        }
      } else {
        int j = side ? my - 1 : 0, jn = side ? my - 2 : 1;
        if ((sb - 8) % 2 == 0) { d = g.at(0, j, q); a = g.at(0, jn, q); b = g.at(1, j, q); }
        else { d = g.at(mx - 1, j, q); a = g.at(mx - 1, jn, q); b = g.at(mx - 2, j, q); }
      }
      u.x[d] = 0.5 * (u.x[a] + u.x[b]);
      u.y[d] = 0.5 * (u.y[a] + u.y[b]);
      u.z[d] = 0.5 * (u.z[a] + u.z[b]);
      if (p) p[d] = 0.5 * (p[a] + p[b]);
    }
    __syncthreads();
  }
}
__global__ void k_edge_average(G g, V3 u, double *p) { edge_average_body(g, u, p); }

// periodic full-plane copies (bcs.c:5567-5731); axis 0 then 2.  x_only:
// the reference copies only .x on the upper x plane in FormBCS.
// Only the ring of the plane (a ghost index on one of the other two axes)
// is copied: periodic_planes_body copied the rest of the plane earlier in
// the same pass (all three components, p and nvert alike), and between the
// two neither those cells nor their sources change (the edge averages write
// points with two ghost indices, the x copy writes i = 0 and mx-1 only), so
// a second copy of the plane's interior rewrites the values it holds.
__device__ __forceinline__ long periodic_ring_size(const G &g, int axis) {
  const int na = axis == 0 ? g.my : g.mx, nb = axis == 0 ? g.mz : g.my;   // plane extents, a fastest
  return 2L * nb + 2L * max(na - 2, 0);
}
__device__ __forceinline__ void periodic_full_body(G g, int axis, int x_only, V3 u, double *p, double *nv, GS gs) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  const int na = axis == 0 ? my : mx, nb = axis == 0 ? mz : my;
  const long n = periodic_ring_size(g, axis);
  for (long t = gs.t0; t < n; t += gs.step) {
    int a, b;   // a = 0 or na-1 over every b, then b = 0 or nb-1 over the inner a
    if (t < 2L * nb) {
      a = (t & 1) ? na - 1 : 0;
      b = (int)(t >> 1);
    } else {
      const long r = t - 2L * nb;
      b = (r & 1) ? nb - 1 : 0;
      a = 1 + (int)(r >> 1);
    }
    long d0, s0, d1, s1;
    if (axis == 0) {
      d0 = g.at(0, a, b); s0 = g.at(mx - 2, a, b); d1 = g.at(mx - 1, a, b); s1 = g.at(1, a, b);
    } else {
      d0 = g.at(a, b, 0); s0 = g.at(a, b, mz - 2); d1 = g.at(a, b, mz - 1); s1 = g.at(a, b, 1);
    }
    u.x[d0] = u.x[s0]; u.y[d0] = u.y[s0]; u.z[d0] = u.z[s0];
    u.x[d1] = u.x[s1];
    if (!(axis == 0 && x_only)) { u.y[d1] = u.y[s1]; u.z[d1] = u.z[s1]; }
    if (p) { p[d0] = p[s0]; p[d1] = p[s1]; nv[d0] = nv[s0]; nv[d1] = nv[s1]; }
  }
}
__global__ void k_periodic_full(G g, int axis, int x_only, V3 u, double *p, double *nv) { periodic_full_body(g, axis, x_only, u, p, nv, GS_SELF); }


// zero wall-normal flux of bctype-1 faces after an RK stage (the build's
// deliberate deviation: the reference never resets it).  bx, by, bz: bit 0
// the lower face, bit 1 the upper face of the axis is a wall
__device__ __forceinline__ void wall_flux_zero_body(G g, int bx, int by, int bz, V3 u, GS gs) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  const long nyz = (long)my * mz, nxz = (long)mx * mz, nxy = (long)mx * my;
  for (long t = gs.t0; t < nyz + nxz + nxy; t += gs.step) {
    if (t < nyz) {
      if (!bx) continue;
      int j = (int)(t % my), k = (int)(t / my);
      if (bx & 1) u.x[g.at(0, j, k)] = 0.;
      if (bx & 2) u.x[g.at(mx - 2, j, k)] = 0.;
    } else if (t < nyz + nxz) {
      if (!by) continue;
      long r = t - nyz;
      int i = (int)(r % mx), k = (int)(r / mx);
      if (by & 1) u.y[g.at(i, 0, k)] = 0.;
      if (by & 2) u.y[g.at(i, my - 2, k)] = 0.;
    } else {
      if (!bz) continue;
      long r = t - nyz - nxz;
      int i = (int)(r % mx), j = (int)(r / mx);
      if (bz & 1) u.z[g.at(i, j, 0)] = 0.;
      if (bz & 2) u.z[g.at(i, j, mz - 2)] = 0.;
    }
  }
}
__global__ void k_wall_flux_zero(G g, int bx, int by, int bz, V3 u) { wall_flux_zero_body(g, bx, by, bz, u, GS_SELF); }

// wall masks of wall_flux_zero_body for one axis from its two bctypes
inline int wall_mask(int lo, int hi) { return (lo == 1 ? 1 : 0) | (hi == 1 ? 2 : 0); }


// ------------------------------------------------------------------
// inlet (bctype 5, k = 0), outlet (bctype 4, k = mz-1), slip (3) and far
// field (6): InflowFlux bcs.c:99-246, OutflowFlux bcs.c:403-469, FormBCS's
// far-field, slip and outlet sections bcs.c:4500-5010.  Kernels over the
// interior points of the boundary faces; sums go through per-block
// partials and a fixed-order final pass.
// ------------------------------------------------------------------
// the analytic inflow profiles (bcs.c:158-196) at the offset face centre
__device__ double inlet_uin(int prof, double xc, double yc, double zc) {
  const double H = 4.1, Umax = 1.5;
  const double r = sqrt(zc * zc + yc * yc);
  double uin = 0.;
  if (prof == 1) uin = 2.;
  else if (prof == -1) uin = -1.;
  else if (prof == 2) uin = 4. * Umax * yc * (H - yc) / (H * H);
  else if (prof == 4) {            // fully developed pipe flow
    uin = 2. * (1. - 4. * r * r);
    if (r > 0.5) uin = 0.;
  } else if (prof == 7) uin = 1.5 * (1. - 4. * yc * yc);
  else if (prof == 10) {           // fully developed flow in a 2 x 2 duct, 20 terms
    const double pi = 3.14159265358979323846264338327950288;
    const double y2 = (yc - 1.5) * 2., x2 = xc - 0.5, A = 1.;
    uin = 1 - y2 * y2;
    for (int n = 0; n < 20; n++)
      uin -= 4. * pow(2. / pi, 3) * pow(-1., n) / pow(2 * n + 1., 3.) * cosh((2 * n + 1) * pi * x2 / 2.) *
             cos((2. * n + 1) * pi * y2 / 2.) / cosh((2 * n + 1) * pi * A / 2.);
  } else if (prof == 11) uin = 0.185;
  return uin;
}

struct Inlet {
  int prof;
  double cx, cy, cz;   // -x_c -y_c -z_c
};

// inlet fluxes and Ubcs where the first cell is fluid; the fluid inflow
// sum to part[block]
__global__ void k_inflow(G g, Inlet in, const double *coor, long n, C3 zet, const double *nv, double *ucz, double *ubf,
                         double *part) {
  __shared__ double sh[BS / 64];
  const long m2 = (long)(g.mx - 2) * (g.my - 2);
  const double *X = coor, *Y = coor + n, *Z = coor + 2 * n;
  double f = 0.0;
  for (long t = blockIdx.x * (long)BS + threadIdx.x; t < m2; t += (long)gridDim.x * BS) {
    const int i = 1 + (int)(t % (g.mx - 2)), j = 1 + (int)(t / (g.mx - 2));
    const long c = g.at(i, j, 0), cs = c - g.mx, cw = c - 1, csw = c - g.mx - 1;   // nodes (i,j) (i,j-1) (i-1,j) (i-1,j-1)
    const double xc = (X[c] + X[cs] + X[cw] + X[csw]) * 0.25 - in.cx;
    const double yc = (Y[c] + Y[cs] + Y[cw] + Y[csw]) * 0.25 - in.cy;
    const double zc = (Z[c] + Z[cs] + Z[cw] + Z[csw]) * 0.25 - in.cz;
    const double uin = inlet_uin(in.prof, xc, yc, zc);
    if (nv[c + g.sk] < 0.1) {
      const double zx = zet.x[c], zy = zet.y[c], zz = zet.z[c];
      const double a = sqrt(zz * zz + zy * zy + zx * zx);
      ubf[ubf_at(g, 4, 1, c)] = uin * zy / a;
      ubf[ubf_at(g, 4, 2, c)] = uin * zz / a;
      ubf[ubf_at(g, 4, 0, c)] = uin * zx / a;
      const double q = uin * a;
      ucz[c] = q;
      f += q;
    }
  }
  const double sf = cv_block_sum(f, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sf;
}

// outlet: velocity flux and area over the last fluid cells (k = mz-2) to
// part[block] and part[nb + block]
__global__ void k_outlet_sum(G g, C3 ucat, C3 zet, const double *nv, double *part) {
  __shared__ double sh[BS / 64];
  const long m2 = (long)(g.mx - 2) * (g.my - 2);
  double f = 0.0, a = 0.0;
  for (long t = blockIdx.x * (long)BS + threadIdx.x; t < m2; t += (long)gridDim.x * BS) {
    const long c = g.at(1 + (int)(t % (g.mx - 2)), 1 + (int)(t / (g.mx - 2)), g.mz - 2);
    if (nv[c] < 0.1) {
      const double zx = zet.x[c], zy = zet.y[c], zz = zet.z[c];
      f += ucat.x[c] * zx + ucat.y[c] * zy + ucat.z[c] * zz;
      a += sqrt(zx * zx + zy * zy + zz * zz);
    }
  }
  const double sf = cv_block_sum(f, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sf;
  const double sa = cv_block_sum(a, sh);
  if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = sa;
}

// outlet correction: Ubcs = Ucat of the last cell, its flux plus the
// uniform correction (FluxInSum - flux) / area; the corrected fluid
// outflow to part[block] when stat
__global__ void k_outlet_fix(G g, C3 ucat, C3 zet, const double *nv, double *ucz, double *ubf, const double *dsc,
                             int stat, double *part) {
  __shared__ double sh[BS / 64];
  const long nxy = (long)g.mx * g.my, m2 = (long)(g.mx - 2) * (g.my - 2), base = (long)(g.mz - 2) * nxy;
  // FluxIn = FluxInSum + FarFluxInSum + FluxIntpSum, the last 0 here
  const double ratio = (dsc[CV_DSC_FLUXIN] + dsc[CV_DSC_FAR] - dsc[CV_DSC_OUTSUM]) / dsc[CV_DSC_OUTSUM + 1];
  double f = 0.0;
  for (long t = blockIdx.x * (long)BS + threadIdx.x; t < m2; t += (long)gridDim.x * BS) {
    const long r = g.at(1 + (int)(t % (g.mx - 2)), 1 + (int)(t / (g.mx - 2)), 0), c = base + r;
    if (nv[c] < 0.1) {
      const double bx = ucat.x[c], by = ucat.y[c], bz = ucat.z[c];
      ubf[ubf_at(g, 5, 0, r)] = bx; ubf[ubf_at(g, 5, 1, r)] = by; ubf[ubf_at(g, 5, 2, r)] = bz;
      const double zx = zet.x[c], zy = zet.y[c], zz = zet.z[c];
      const double q = (bx * zx + by * zy + bz * zz) + ratio * sqrt(zx * zx + zy * zy + zz * zz);
      ucz[c] = q;
      f += q;
    }
  }
  if (!stat) return;
  const double sf = cv_block_sum(f, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sf;
}

// point t of the interior points of the six faces, in face order (i-, i+,
// j-, j+, k-, k+; each face's points i fastest, then j, then k): the face
// f, the storage indices of the ghost cell, the adjacent cell and the
// boundary face (0 on lower faces, m-2 on upper ones), and the ghost's
// plane index r.  Returns false past the last point.
struct FacePt {
  int f;
  long gi, ai, bi, r;
};
__device__ inline bool face_point(const G &g, long t, FacePt &p) {
  const long nx = g.mx - 2, ny = g.my - 2, nz = g.mz - 2;
  const long sz[3] = {ny * nz, nx * nz, nx * ny};
  int f = 0;
  while (f < 6 && t >= sz[f / 2]) { t -= sz[f / 2]; f++; }
  if (f == 6) return false;
  const int d = f / 2, up = f & 1;
  const int m = d == 0 ? g.mx : (d == 1 ? g.my : g.mz);
  const int gg = up ? m - 1 : 0, aa = up ? m - 2 : 1, bb = up ? m - 2 : 0;
  if (d == 0) {
    const int j = 1 + (int)(t % ny), k = 1 + (int)(t / ny);
    p.gi = g.at(gg, j, k); p.ai = g.at(aa, j, k); p.bi = g.at(bb, j, k); p.r = j + (long)g.my * k;
  } else if (d == 1) {
    const int i = 1 + (int)(t % nx), k = 1 + (int)(t / nx);
    p.gi = g.at(i, gg, k); p.ai = g.at(i, aa, k); p.bi = g.at(i, bb, k); p.r = i + (long)g.mx * k;
  } else {
    const int i = 1 + (int)(t % nx), j = 1 + (int)(t / nx);
    p.gi = g.at(i, j, gg); p.ai = g.at(i, j, aa); p.bi = g.at(i, j, bb); p.r = i + (long)g.mx * j;
  }
  p.f = f;
  return true;
}
__host__ __device__ inline long face_points(int mx, int my, int mz) {
  return 2 * ((long)(my - 2) * (mz - 2) + (long)(mx - 2) * (mz - 2) + (long)(mx - 2) * (my - 2));
}

// slip faces (bctype 3; FormBCS bcs.c:4729-4853, GhostNodeVelocity
// bcs.c:3833-3960): Ubcs = the adjacent Ucat less its normal component;
// with formbcs the face's normal flux is zeroed
__global__ void k_slip(G g, int mask, int formbcs, C3 ucat, C3 csi, C3 eta, C3 zet, V3 ucont, double *ubf) {
  const long np = face_points(g.mx, g.my, g.mz);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < np; t += (long)gridDim.x * blockDim.x) {
    FacePt p;
    if (!face_point(g, t, p) || !((mask >> p.f) & 1)) continue;
    const int d = p.f / 2;
    const C3 m = d == 0 ? csi : (d == 1 ? eta : zet);
    const double mx_ = m.x[p.bi], my_ = m.y[p.bi], mz_ = m.z[p.bi];
    const double A = sqrt(mz_ * mz_ + my_ * my_ + mx_ * mx_);
    const double nx = mx_ / A, ny = my_ / A, nz = mz_ / A;
    const double ux = ucat.x[p.ai], uy = ucat.y[p.ai], uz = ucat.z[p.ai];
    const double Un = ux * nx + uy * ny + uz * nz;
    ubf[ubf_at(g, p.f, 0, p.r)] = ux - Un * nx;
    ubf[ubf_at(g, p.f, 1, p.r)] = uy - Un * ny;
    ubf[ubf_at(g, p.f, 2, p.r)] = uz - Un * nz;
    if (formbcs) (d == 0 ? ucont.x : (d == 1 ? ucont.y : ucont.z))[p.bi] = 0.;
  }
}

// far-field faces (bctype 6; FormBCS bcs.c:4500-4608): Ubcs = the adjacent
// Ucat and the face flux its normal component times the matching metric
// component; the inflow (lower faces), outflow (upper faces), |flux| and
// the two areas to part[q*nb + block], q = 0..4
__global__ void k_far_sum(G g, int mask, C3 ucat, C3 csi, C3 eta, C3 zet, V3 ucont, double *ubf, double *part) {
  __shared__ double sh[BS / 64];
  const long np = face_points(g.mx, g.my, g.mz);
  double fin = 0.0, fout = 0.0, fab = 0.0, ain = 0.0, aout = 0.0;
  for (long t = blockIdx.x * (long)BS + threadIdx.x; t < np; t += (long)gridDim.x * BS) {
    FacePt p;
    if (!face_point(g, t, p) || !((mask >> p.f) & 1)) continue;
    const int d = p.f / 2;
    const double ux = ucat.x[p.ai], uy = ucat.y[p.ai], uz = ucat.z[p.ai];
    ubf[ubf_at(g, p.f, 0, p.r)] = ux;
    ubf[ubf_at(g, p.f, 1, p.r)] = uy;
    ubf[ubf_at(g, p.f, 2, p.r)] = uz;
    const double mc = d == 0 ? csi.x[p.bi] : (d == 1 ? eta.y[p.bi] : zet.z[p.bi]);
    const double q = (d == 0 ? ux : (d == 1 ? uy : uz)) * mc;
    (d == 0 ? ucont.x : (d == 1 ? ucont.y : ucont.z))[p.bi] = q;
    if (p.f & 1) { fout += q; aout += mc; }
    else { fin += q; ain += mc; }
    fab += fabs(q);
  }
  const double v[5] = {fin, fout, fab, ain, aout};
  for (int q = 0; q < 5; q++) {
    const double sv = cv_block_sum(v[q], sh);
    if (threadIdx.x == 0) part[q * gridDim.x + blockIdx.x] = sv;
  }
}

// far-field flux balance (bcs.c:4610-4724): with a far-field upper face,
// ratio = (inflow - outflow) / sum |flux| (0 below 1e-10); upper faces
// gain and lower faces lose ratio * |flux|, and Ubcs's normal component
// follows the corrected flux
__global__ void k_far_fix(G g, int mask, int has_out, C3 csi, C3 eta, C3 zet, V3 ucont, double *ubf,
                          const double *dsc) {
  double ratio = 0.;
  if (has_out) {
    const double fab = dsc[CV_DSC_FAR + 2];
    ratio = (dsc[CV_DSC_FAR] - dsc[CV_DSC_FAR + 1]) / fab;
    if (fabs(fab) < 1.e-10) ratio = 0.;
  }
  const long np = face_points(g.mx, g.my, g.mz);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < np; t += (long)gridDim.x * blockDim.x) {
    FacePt p;
    if (!face_point(g, t, p) || !((mask >> p.f) & 1)) continue;
    const int d = p.f / 2;
    double *uq = d == 0 ? ucont.x : (d == 1 ? ucont.y : ucont.z);
    double q = uq[p.bi];
    if (p.f & 1) q += ratio * fabs(q);
    else q -= ratio * fabs(q);
    uq[p.bi] = q;
    const double mc = d == 0 ? csi.x[p.bi] : (d == 1 ? eta.y[p.bi] : zet.z[p.bi]);
    ubf[ubf_at(g, p.f, d, p.r)] = q / mc;
  }
}

// ------------------------------------------------------------------
// ibm_interpolation_advanced (ibm.c:3012-3254) with Add_dUndt, no wall
// function: per IB node the velocity of the interception point (a
// weighted sum of three cells), the no-slip (wallfunction.c:12-27) or
// free-slip (wallfunction.c:29-46) reconstruction at the node, and p from
// the same weights plus the dU_n/dt term.  The new values go to ib_new
// first: every node reads the Ucat and p of the sweep's start, as the
// reference reads lUcat and lP.
// ------------------------------------------------------------------
__device__ __forceinline__ void ib_node_body(int q, const int *node, const double *w, int freeslip, double dt, C3 ucat,
                                             const double *p, int with_p, double *out) {
  const int *I = node + 4 * q;
  const double *W = w + 12 * q;
  const long a1 = I[1], a2 = I[2], a3 = I[3];
  const double sk1 = W[0], sk2 = W[1], sk3 = W[2];
  const double sb = W[3], sc = sb + W[4];
  const double uax = W[5], uay = W[6], uaz = W[7];
  const double ucx = (ucat.x[a1] * sk1 + ucat.x[a2] * sk2 + ucat.x[a3] * sk3);
  const double ucy = (ucat.y[a1] * sk1 + ucat.y[a2] * sk2 + ucat.y[a3] * sk3);
  const double ucz = (ucat.z[a1] * sk1 + ucat.z[a2] * sk2 + ucat.z[a3] * sk3);
  double bx, by, bz;
  if (freeslip) {
    const double nx = W[9], ny = W[10], nz = W[11];
    const double Uan = uax * nx + uay * ny + uaz * nz;
    const double Ucn = ucx * nx + ucy * ny + ucz * nz;
    const double Ubn = Uan + (Ucn - Uan) * sb / sc;
    bx = ucx - Ucn * nx; by = ucy - Ucn * ny; bz = ucz - Ucn * nz;
    bx += Ubn * nx; by += Ubn * ny; bz += Ubn * nz;
  } else {
    bx = sb / sc * (ucx - uax); by = sb / sc * (ucy - uay); bz = sb / sc * (ucz - uaz);
    bx += uax; by += uay; bz += uaz;
  }
  double *o = out + 4 * q;
  o[0] = bx; o[1] = by; o[2] = bz;
  if (with_p) {
    double pn = (p[a1] * sk1 + p[a2] * sk2 + p[a3] * sk3);
    pn += W[8] / dt * (sc - sb);
    o[3] = pn;
  }
}
__global__ void k_ib_nodes(int n, const int *node, const double *w, int freeslip, double dt, C3 ucat, const double *p,
                           int with_p, double *out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) ib_node_body(q, node, w, freeslip, dt, ucat, p, with_p, out);
}

__device__ __forceinline__ void ib_scatter_body(int q, const int *node, const double *in, V3 ucat, double *p, int with_p) {
  const long t = node[4 * q];
  const double *v = in + 4 * q;
  ucat.x[t] = v[0]; ucat.y[t] = v[1]; ucat.z[t] = v[2];
  if (with_p) p[t] = v[3];
}
__global__ void k_ib_scatter(int n, const int *node, const double *in, V3 ucat, double *p, int with_p) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) ib_scatter_body(q, node, in, ucat, p, with_p);
}

// the fluxes of the faces touching an IB node, from the average of the
// two cells' Ucat (ibm.c:3166-3224); cells listed x faces, then y, then z
__device__ __forceinline__ void ib_face_body(int q, G g, int nx, int ny, const int *cell, C3 ucat, C3 csi, C3 eta,
                                             C3 zet, V3 ucont) {
  const long t = cell[q];
  if (q < nx) {
    const long u = t + 1;
    const double ux = (ucat.x[t] + ucat.x[u]) * 0.5, uy = (ucat.y[t] + ucat.y[u]) * 0.5, uz = (ucat.z[t] + ucat.z[u]) * 0.5;
    ucont.x[t] = ux * csi.x[t] + uy * csi.y[t] + uz * csi.z[t];
  } else if (q < nx + ny) {
    const long u = t + g.sj;
    const double ux = (ucat.x[u] + ucat.x[t]) * 0.5, uy = (ucat.y[u] + ucat.y[t]) * 0.5, uz = (ucat.z[u] + ucat.z[t]) * 0.5;
    ucont.y[t] = ux * eta.x[t] + uy * eta.y[t] + uz * eta.z[t];
  } else {
    const long u = t + g.sk;
    const double ux = (ucat.x[u] + ucat.x[t]) * 0.5, uy = (ucat.y[u] + ucat.y[t]) * 0.5, uz = (ucat.z[u] + ucat.z[t]) * 0.5;
    ucont.z[t] = ux * zet.x[t] + uy * zet.y[t] + uz * zet.z[t];
  }
}
__global__ void k_ib_faces(G g, int nx, int ny, int nz, const int *cell, C3 ucat, C3 csi, C3 eta, C3 zet, V3 ucont) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nx + ny + nz) ib_face_body(q, g, nx, ny, cell, ucat, csi, eta, zet, ucont);
}

// Contra2Cart of the listed interior cells
__global__ void k_ib_c2c(G g, int n, const int *cell, C3 uc, C3 csi, C3 eta, C3 zet, const double *nv, V3 ucat) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int i, j, k;
  g.dec(cell[q], i, j, k);
  c2c_cell(g, uc, csi, eta, zet, nv, ucat, i, j, k);
}

// OutflowFlux: the k = mz-2 face flux over the interior, fluid or not
__global__ void k_outflow_diag(G g, const double *ucz, double *part) {
  __shared__ double sh[BS / 64];
  const long m2 = (long)(g.mx - 2) * (g.my - 2);
  double f = 0.0;
  for (long t = blockIdx.x * (long)BS + threadIdx.x; t < m2; t += (long)gridDim.x * BS)
    f += ucz[g.at(1 + (int)(t % (g.mx - 2)), 1 + (int)(t / (g.mx - 2)), g.mz - 2)];
  const double sf = cv_block_sum(f, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sf;
}


// ------------------------------------------------------------------
// Face fluxes: QUICK convection (rhs.c:1636-1897) and the viscous flux
// (rhs.c:2062-2486).  The flux rules are written once over an accessor
// that returns ucat and the nvert classes at offsets from the face's cell,
// so the same arithmetic runs on global loads and on LDS tiles.
// ------------------------------------------------------------------
// nvert classes the face rules test
#define NVB_LT01 1u   // nv < 0.1
#define NVB_GT01 2u   // nv > 0.1
#define NVB_GT7 4u    // nv > 7 (inner blank)
#define NVB_SOL 8u    // 0.5 < nv < 7 (solid)
__device__ inline unsigned nv_bits(double v) {
  return (v < 0.1 ? NVB_LT01 : 0u) | (v > 0.1 ? NVB_GT01 : 0u) | (v > 7. ? NVB_GT7 : 0u) |
         ((v > 0.5 && v < 7.) ? NVB_SOL : 0u);
}

// unit offset along axis a, scaled by s
#define EOFS(a, s) ((a) == 0 ? (s) : 0), ((a) == 1 ? (s) : 0), ((a) == 2 ? (s) : 0)

// ucat and nvert at offsets of the face's cell, from global memory
struct GAcc {
  const G *g;
  C3 u;
  const double *nv;
  int i, j, k;
  __device__ long idx(int di, int dj, int dk) const { return g->atq(i + di, j + dj, k + dk); }
  __device__ void u3(int di, int dj, int dk, double *v) const {
    const long p = idx(di, dj, dk);
    v[0] = u.x[p]; v[1] = u.y[p]; v[2] = u.z[p];
  }
  __device__ unsigned nb(int di, int dj, int dk) const { return nv_bits(nv[idx(di, dj, dk)]); }
};

// the same from an LDS tile of planes k-1, k, k+1 (row pitch TW); the
// k+2 values of the face's own column come in registers
template <int TW>
struct TAcc {
  const double *P[3][3];          // [plane k-1, k, k+1][x, y, z]
  const unsigned char *N[3];      // nvert classes of the planes
  int o;                          // tile offset of the face's cell
  double kp2[3];                  // ucat at (i, j, k+2)
  __device__ void u3(int di, int dj, int dk, double *v) const {
    if (dk == 2) { v[0] = kp2[0]; v[1] = kp2[1]; v[2] = kp2[2]; return; }
    const int e = o + dj * TW + di;
    v[0] = P[dk + 1][0][e]; v[1] = P[dk + 1][1][e]; v[2] = P[dk + 1][2][e];
  }
  __device__ unsigned nb(int di, int dj, int dk) const { return N[dk + 1][o + dj * TW + di]; }
};

__device__ inline void quick3(double um, double up, const double *w2, const double *v1, const double *v0,
                              const double *wm, double *f) {
  const double coef = 0.125;
#pragma unroll
  for (int c = 0; c < 3; c++)
    f[c] = um * (coef * (-w2[c] - 2. * v1[c] + 3. * v0[c]) + v1[c]) + up * (coef * (-wm[c] - 2. * v0[c] + 3. * v1[c]) + v0[c]);
}

// convective flux through face (i,j,k) of axis D; ucd is the face's
// contravariant flux.  The far upwind points are +2 (or +1 / the cell
// itself near walls and blanked cells) and -1 (or the cell itself).
template <int D, class Acc>
__device__ inline void conv_acc(const G &g, int i, int j, int k, double ucd, const Acc &a, double *f) {
  const double ucon = ucd * 0.5, up = ucon + fabs(ucon), um = ucon - fabs(ucon);
  const int c = D == 0 ? i : D == 1 ? j : k;
  const int m = D == 0 ? g.mx : D == 1 ? g.my : g.mz;
  const bool per = D == 0 ? g.xper != 0 : D == 2 ? g.zper != 0 : false;
  const unsigned np1 = a.nb(EOFS(D, 1));
  const unsigned nm1 = (c > 0 || per) ? a.nb(EOFS(D, -1)) : NVB_LT01;
  const bool fr1 = (np1 & (NVB_LT01 | NVB_GT7)) != 0, frm = (nm1 & (NVB_LT01 | NVB_GT7)) != 0;
  int s2, sm;   // far points: s2 in {2, 1, 0}, sm in {-1, 0}
  if (D == 1) {
    const int A2 = j + 2 < g.my ? 2 : 0, AM = j > 0 ? -1 : 0;
    if (j > 0 && j < g.my - 2 && fr1 && frm) { s2 = A2; sm = AM; }
    else if (j == 0 || (nm1 & NVB_GT01)) { s2 = A2; sm = 0; }
    else if (j == g.my - 2 || (np1 & NVB_GT01)) { s2 = 1; sm = AM; }
    else { f[0] = f[1] = f[2] = 0.0; return; }
  } else {
    const bool both = (nm1 & NVB_LT01) && (np1 & NVB_LT01);
    if (c > 0 && c < m - 2 && fr1 && frm) { s2 = 2; sm = -1; }
    else if ((D == 0 || c < m - 2) && (c == 0 || (nm1 & NVB_GT01))) {
      s2 = 2; sm = (per && c == 0 && both) ? -1 : 0;
    } else if ((D == 0 || c > 0) && (c == m - 2 || (np1 & NVB_GT01))) {
      const bool wrap = per && c == m - 2 && both;
      s2 = wrap ? 2 : 1; sm = -1;
    } else { f[0] = f[1] = f[2] = 0.0; return; }
  }
  double v1[3], v0[3], w2[3], wm[3];
  a.u3(EOFS(D, 1), v1);
  a.u3(0, 0, 0, v0);
  if (s2 == 2) a.u3(EOFS(D, 2), w2);
  else for (int q = 0; q < 3; q++) w2[q] = s2 == 1 ? v1[q] : v0[q];
  if (sm == -1) a.u3(EOFS(D, -1), wm);
  else for (int q = 0; q < 3; q++) wm[q] = v0[q];
  quick3(um, up, w2, v1, v0, wm, f);
}

struct FM { C3 c, e, z; const double *aj; };

// viscous flux through face t of axis D; tangential differences switch to
// one-sided next to solid cells
template <int D, class Acc>
__device__ inline void visc_acc(long t, const Acc &a, const FM &m, double nu, double *f) {
  double du[3][3], ut[3], uu[3];
  a.u3(0, 0, 0, ut);
  a.u3(EOFS(D, 1), uu);
#pragma unroll
  for (int tt = 0; tt < 3; tt++) {
    if (tt == D) {
      for (int q = 0; q < 3; q++) du[q][tt] = uu[q] - ut[q];
      continue;
    }
    const unsigned n1 = a.nb(EOFS(tt, 1)), n3 = a.nb(EOFS(tt, -1));
    const unsigned n2 = a.nb((D == 0) + (tt == 0), (D == 1) + (tt == 1), (D == 2) + (tt == 2));
    const unsigned n4 = a.nb((D == 0) - (tt == 0), (D == 1) - (tt == 1), (D == 2) - (tt == 2));
    const bool hi = ((n1 | n2) & NVB_SOL) != 0, lo = ((n3 | n4) & NVB_SOL) != 0;
    double tp[3], tm[3], pp[3], pm[3];   // t+s, t-s, up+s, up-s
    if (!hi) {
      a.u3(EOFS(tt, 1), tp);
      a.u3((D == 0) + (tt == 0), (D == 1) + (tt == 1), (D == 2) + (tt == 2), pp);
    }
    if (hi || !lo) {
      a.u3(EOFS(tt, -1), tm);
      a.u3((D == 0) - (tt == 0), (D == 1) - (tt == 1), (D == 2) - (tt == 2), pm);
    }
    for (int q = 0; q < 3; q++) {
      if (hi) du[q][tt] = (uu[q] + ut[q] - pm[q] - tm[q]) * 0.5;
      else if (lo) du[q][tt] = (pp[q] + tp[q] - uu[q] - ut[q]) * 0.5;
      else du[q][tt] = (pp[q] + tp[q] - pm[q] - tm[q]) * 0.25;
    }
  }
  const double csi0 = m.c.x[t], csi1 = m.c.y[t], csi2 = m.c.z[t];
  const double eta0 = m.e.x[t], eta1 = m.e.y[t], eta2 = m.e.z[t];
  const double zet0 = m.z.x[t], zet1 = m.z.y[t], zet2 = m.z.z[t];
  double n0, n1, n2, g11, g21, g31;
  if (D == 0) {
    n0 = csi0; n1 = csi1; n2 = csi2;
    g11 = csi0 * csi0 + csi1 * csi1 + csi2 * csi2; g21 = eta0 * csi0 + eta1 * csi1 + eta2 * csi2;
    g31 = zet0 * csi0 + zet1 * csi1 + zet2 * csi2;
  } else if (D == 1) {
    n0 = eta0; n1 = eta1; n2 = eta2;
    g11 = csi0 * eta0 + csi1 * eta1 + csi2 * eta2; g21 = eta0 * eta0 + eta1 * eta1 + eta2 * eta2;
    g31 = zet0 * eta0 + zet1 * eta1 + zet2 * eta2;
  } else {
    n0 = zet0; n1 = zet1; n2 = zet2;
    g11 = csi0 * zet0 + csi1 * zet1 + csi2 * zet2; g21 = eta0 * zet0 + eta1 * zet1 + eta2 * zet2;
    g31 = zet0 * zet0 + zet1 * zet1 + zet2 * zet2;
  }
  const double dudc = du[0][0], dude = du[0][1], dudz = du[0][2];
  const double dvdc = du[1][0], dvde = du[1][1], dvdz = du[1][2];
  const double dwdc = du[2][0], dwde = du[2][1], dwdz = du[2][2];
  const double r11 = dudc * csi0 + dude * eta0 + dudz * zet0;
  const double r21 = dvdc * csi0 + dvde * eta0 + dvdz * zet0;
  const double r31 = dwdc * csi0 + dwde * eta0 + dwdz * zet0;
  const double r12 = dudc * csi1 + dude * eta1 + dudz * zet1;
  const double r22 = dvdc * csi1 + dvde * eta1 + dvdz * zet1;
  const double r32 = dwdc * csi1 + dwde * eta1 + dwdz * zet1;
  const double r13 = dudc * csi2 + dude * eta2 + dudz * zet2;
  const double r23 = dvdc * csi2 + dvde * eta2 + dvdz * zet2;
  const double r33 = dwdc * csi2 + dwde * eta2 + dwdz * zet2;
  const double ajc = m.aj[t];
  f[0] = 0; f[1] = 0; f[2] = 0;
  f[0] += (g11 * dudc + g21 * dude + g31 * dudz + r11 * n0 + r21 * n1 + r31 * n2) * ajc * (nu);
  f[1] += (g11 * dvdc + g21 * dvde + g31 * dvdz + r12 * n0 + r22 * n1 + r32 * n2) * ajc * (nu);
  f[2] += (g11 * dwdc + g21 * dwde + g31 * dwdz + r13 * n0 + r23 * n1 + r33 * n2) * ajc * (nu);
}

// Visc - Conv flux through face (i,j,k) of family D (the faces
// rhs.c:1636-1897 and rhs.c:2062-2486 evaluate into Fp1/Fp2/Fp3)
template <int D, class Acc>
__device__ inline void face_rc(const G &g, int i, int j, int k, long t, double ucd, const Acc &a, const FM &m, double nu,
                               double *F) {
  double cf[3], vf[3];
  conv_acc<D>(g, i, j, k, ucd, a, cf);
  visc_acc<D>(t, a, m, nu, vf);
  F[0] = vf[0] - cf[0];
  F[1] = vf[1] - cf[1];
  F[2] = vf[2] - cf[2];
}

// rct = aj*(avg basis).Rc of interior cell t from the face-flux
// differences along i, j and k (rhs.c:2663-2703)
__device__ inline void rct_out(const G &g, long t, const double *d0, const double *d1, const double *d2, C3 csi, C3 eta,
                               C3 zet, const double *aj, V3 rct) {
  const long ti = t - 1, tj = t - g.sj, tk = t - g.sk;
  const double rcx = d0[0] + d1[0] + d2[0];
  const double rcy = d0[1] + d1[1] + d2[1];
  const double rcz = d0[2] + d1[2] + d2[2];
  const double A = aj[t];
  rct.x[t] = A * (0.5 * (csi.x[t] + csi.x[ti]) * rcx + 0.5 * (csi.y[t] + csi.y[ti]) * rcy + 0.5 * (csi.z[t] + csi.z[ti]) * rcz);
  rct.y[t] = A * (0.5 * (eta.x[t] + eta.x[tj]) * rcx + 0.5 * (eta.y[t] + eta.y[tj]) * rcy + 0.5 * (eta.z[t] + eta.z[tj]) * rcz);
  rct.z[t] = A * (0.5 * (zet.x[t] + zet.x[tk]) * rcx + 0.5 * (zet.y[t] + zet.y[tk]) * rcy + 0.5 * (zet.z[t] + zet.z[tk]) * rcz);
}

// rct_out in two halves: the face-averaged basis vectors and aj of cell t
// (loaded ahead, e.g. before a barrier), then the products with Rc.  The
// terms and their order are rct_out's, so the bits are the same
struct RctOps {
  double A, c[3], e[3], z[3];
};
// zt, zb: zet at t and at t - sk (the k-marching pass carries the latter)
__device__ inline RctOps rct_load(const G &g, long t, C3 csi, C3 eta, const double *zt, const double *zb,
                                  const double *aj) {
  const long ti = t - 1, tj = t - g.sj;
  RctOps o;
  o.A = aj[t];
  o.c[0] = 0.5 * (csi.x[t] + csi.x[ti]); o.c[1] = 0.5 * (csi.y[t] + csi.y[ti]); o.c[2] = 0.5 * (csi.z[t] + csi.z[ti]);
  o.e[0] = 0.5 * (eta.x[t] + eta.x[tj]); o.e[1] = 0.5 * (eta.y[t] + eta.y[tj]); o.e[2] = 0.5 * (eta.z[t] + eta.z[tj]);
  o.z[0] = 0.5 * (zt[0] + zb[0]); o.z[1] = 0.5 * (zt[1] + zb[1]); o.z[2] = 0.5 * (zt[2] + zb[2]);
  return o;
}
__device__ inline void rct_store(const RctOps &o, long t, const double *d0, const double *d1, const double *d2, V3 rct) {
  const double rcx = d0[0] + d1[0] + d2[0];
  const double rcy = d0[1] + d1[1] + d2[1];
  const double rcz = d0[2] + d1[2] + d2[2];
  rct.x[t] = o.A * (o.c[0] * rcx + o.c[1] * rcy + o.c[2] * rcz);
  rct.y[t] = o.A * (o.e[0] * rcx + o.e[1] * rcy + o.e[2] * rcz);
  rct.z[t] = o.A * (o.z[0] * rcx + o.z[1] * rcy + o.z[2] * rcz);
}

// Rc = Visc - Conv as the divergence of the three face families and rct
// at the interior cells, in one pass with no face arrays in HBM.  Block
// 64 x TY owns an (i, j) tile and marches RC_KC planes of k: lane 0 is the
// halo column i0-1 (its i-face reaches lane 1 through a lane shuffle), row
// 0 the halo row j0-1 (its j-face reaches row 1 through LDS), and the
// k-face of plane k-1 stays in registers.  63 x (TY-1) cells per block and
// plane.  This variant gathers ucat and nvert from global memory.
#define RC_KC 16
template <int TY>
__global__ void __launch_bounds__(64 * TY) k_rc_fused(G g, C3 uc, C3 u, const double *nv, FM mi, FM mj, FM mk, double nu,
                                                    C3 csi, C3 eta, C3 zet, const double *aj, V3 rct, int kc) {
  __shared__ double S[3][TY][64];
  const int L = threadIdx.x, R = threadIdx.y;
  const int mx = g.mx, my = g.my, mz = g.mz;
  const int i = 63 * blockIdx.x - 1 + L, j = (TY - 1) * blockIdx.y - 1 + R;
  const int k0 = kc * blockIdx.z, k1 = min(k0 + kc, mz);
  const bool in_i = i >= 1 && i <= mx - 2, in_j = j >= 1 && j <= my - 2;
  const bool out = L >= 1 && R >= 1 && in_i && in_j;
  const bool do0 = R >= 1 && in_j && i >= 0 && i <= mx - 2;   // i-face (i,j,k)
  const bool do1 = L >= 1 && in_i && j >= 0 && j <= my - 2;   // j-face (i,j,k)
  double p2[3] = {0., 0., 0.};                               // k-face (i,j,k-1)
  if (out && k0 >= 1) {
    const long t = g.at(i, j, k0 - 1);
    GAcc a{&g, u, nv, i, j, k0 - 1};
    face_rc<2>(g, i, j, k0 - 1, t, uc.z[t], a, mk, nu, p2);
  }
  for (int k = k0; k < k1; k++) {
    const bool kin = k >= 1 && k <= mz - 2;
    const long t = g.at(i, j, k);
    const GAcc a{&g, u, nv, i, j, k};
    double d0[3], a1[3] = {0., 0., 0.}, a2[3] = {0., 0., 0.};
    {
      double a0[3] = {0., 0., 0.};
      if (do0 && kin) face_rc<0>(g, i, j, k, t, uc.x[t], a, mi, nu, a0);
#pragma unroll
      for (int c = 0; c < 3; c++) d0[c] = a0[c] - __shfl_up(a0[c], 1, 64);
    }
    if (do1 && kin) face_rc<1>(g, i, j, k, t, uc.y[t], a, mj, nu, a1);
#pragma unroll
    for (int c = 0; c < 3; c++) S[c][R][L] = a1[c];
    if (out && k <= mz - 2) face_rc<2>(g, i, j, k, t, uc.z[t], a, mk, nu, a2);
    __syncthreads();
    if (out && kin) {
      double d1[3], d2[3];
#pragma unroll
      for (int c = 0; c < 3; c++) { d1[c] = a1[c] - S[c][R - 1][L]; d2[c] = a2[c] - p2[c]; }
      rct_out(g, t, d0, d1, d2, csi, eta, zet, aj, rct);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; c++) p2[c] = a2[c];
  }
}

// The same pass with ucat and the nvert classes staged through LDS: a ring
// of three k-planes of the tile i0-2 .. i0+64 x j0-2 .. j0+TY (every point
// the faces of the block read in those planes), one plane loaded per step
// with coalesced rows.  Barrier A publishes the new plane; barrier B
// publishes the j-faces and ends every read of the oldest plane before the
// next step overwrites its slot.  Points outside the grid are loaded
// through the periodic ghost map on x and z or left zero: no face rule
// reads them.
template <int TY>
__device__ __forceinline__ void rc_tiled_body(const G &g, C3 uc, C3 u, const double *nv, const FM &mi, const FM &mj,
                                              const FM &mk, double nu, C3 csi, C3 eta, C3 zet, const double *aj, V3 rct,
                                              int kc, int bx, int by, int bz) {
  constexpr int TW = 67, TH = TY + 3, NP = TW * TH, NT = 64 * TY;
  __shared__ double SU[3][3][NP];
  __shared__ unsigned char SN[3][NP];
  __shared__ double SF[3][TY][64];
  const int L = threadIdx.x, R = threadIdx.y, tid = R * 64 + L;
  const int mx = g.mx, my = g.my, mz = g.mz;
  const int i0 = 63 * bx, j0 = (TY - 1) * by;
  const int i = i0 - 1 + L, j = j0 - 1 + R;
  const int k0 = kc * bz, k1 = min(k0 + kc, mz);
  const bool in_i = i >= 1 && i <= mx - 2, in_j = j >= 1 && j <= my - 2;
  const bool out = L >= 1 && R >= 1 && in_i && in_j;
  const bool do0 = R >= 1 && in_j && i >= 0 && i <= mx - 2;
  const bool do1 = L >= 1 && in_i && j >= 0 && j <= my - 2;
  const int ks = k0 > 0 ? k0 - 1 : 0;   // the k-face of plane k0-1 starts the k difference
  auto load_plane = [&](int kk) {
    const int sl = (kk + 3) % 3;
    int kq = kk;
    bool kok = kk >= 0 && kk < mz;
    if (!kok && g.zper && (kk == -1 || kk == mz)) { kq = kk == -1 ? mz - 3 : 2; kok = true; }
    for (int e = tid; e < NP; e += NT) {
      const int row = e / TW, col = e - row * TW;
      int gi = i0 - 2 + col;
      const int gj = j0 - 2 + row;
      bool ok = kok && gj >= 0 && gj < my;
      if (gi < 0 || gi >= mx) {
        if (g.xper && (gi == -1 || gi == mx)) gi = gi == -1 ? mx - 3 : 2;
        else ok = false;
      }
      double x = 0., y = 0., z = 0., w = 0.;
      if (ok) {
        const long p = g.at(gi, gj, kq);
        x = u.x[p]; y = u.y[p]; z = u.z[p]; w = nv[p];
      }
      SU[sl][0][e] = x; SU[sl][1][e] = y; SU[sl][2][e] = z;
      SN[sl][e] = (unsigned char)nv_bits(w);
    }
  };
  load_plane(ks - 1);
  load_plane(ks);
  double p2[3] = {0., 0., 0.};
  // zet of the previous plane's cell, carried: rct's k average reads it, and
  // a plane of the XCD's blocks outgrows its L2 between the two reads
  double zb[3] = {0., 0., 0.};
  for (int k = ks; k < k1; k++) {
    load_plane(k + 1);
    __syncthreads();   // A
    const bool kin = k >= k0 && k >= 1 && k <= mz - 2;
    const long t = g.at(i, j, k);
    TAcc<TW> a;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int sl = (k - 1 + q + 3) % 3;
      a.P[q][0] = SU[sl][0]; a.P[q][1] = SU[sl][1]; a.P[q][2] = SU[sl][2];
      a.N[q] = SN[sl];
    }
    a.o = (R + 1) * TW + L + 1;
    a.kp2[0] = a.kp2[1] = a.kp2[2] = 0.;
    double d0[3], a1[3] = {0., 0., 0.}, a2[3] = {0., 0., 0.};
    {
      double a0[3] = {0., 0., 0.};
      if (do0 && kin) face_rc<0>(g, i, j, k, t, uc.x[t], a, mi, nu, a0);
#pragma unroll
      for (int c = 0; c < 3; c++) d0[c] = a0[c] - __shfl_up(a0[c], 1, 64);
    }
    if (do1 && kin) face_rc<1>(g, i, j, k, t, uc.y[t], a, mj, nu, a1);
#pragma unroll
    for (int c = 0; c < 3; c++) SF[c][R][L] = a1[c];
    if (out && k <= mz - 2) {
      const int kk = k + 2;
      if (kk < mz || (g.zper && kk == mz)) {
        const long p = g.at(i, j, kk < mz ? kk : 2);
        a.kp2[0] = u.x[p]; a.kp2[1] = u.y[p]; a.kp2[2] = u.z[p];
      }
      face_rc<2>(g, i, j, k, t, uc.z[t], a, mk, nu, a2);
    }
    // rct's operands are loaded before barrier B, so their latency overlaps the wait
    RctOps ro{};
    double zt[3] = {0., 0., 0.};
    if (out) { zt[0] = zet.x[t]; zt[1] = zet.y[t]; zt[2] = zet.z[t]; }
    if (out && kin) ro = rct_load(g, t, csi, eta, zt, zb, aj);
#pragma unroll
    for (int c = 0; c < 3; c++) zb[c] = zt[c];
    __syncthreads();   // B
    if (out && kin) {
      double d1[3], d2[3];
#pragma unroll
      for (int c = 0; c < 3; c++) { d1[c] = a1[c] - SF[c][R - 1][L]; d2[c] = a2[c] - p2[c]; }
      rct_store(ro, t, d0, d1, d2, rct);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) p2[c] = a2[c];
  }
}

template <int TY>
__global__ void __launch_bounds__(64 * TY) k_rc_tiled(G g, C3 uc, C3 u, const double *nv, FM mi, FM mj, FM mk, double nu,
                                                    C3 csi, C3 eta, C3 zet, const double *aj, V3 rct, int kc) {
  rc_tiled_body<TY>(g, uc, u, nv, mi, mj, mk, nu, csi, eta, zet, aj, rct, kc, blockIdx.x, blockIdx.y, blockIdx.z);
}

// periodic copies of rct planes (DMLocalToLocal + rhs.c:2725-2781)
__device__ __forceinline__ void rct_periodic_body(G g, V3 rct, GS gs) {
  const int mx = g.mx, my = g.my, mz = g.mz;
  const long nyz = (long)my * mz, nxy = (long)mx * my;
  for (long t = gs.t0; t < nyz + nxy; t += gs.step) {
    if (t < nyz) {
      if (!g.xper) continue;
      int j = (int)(t % my), k = (int)(t / my);
      if (!(j > 0 && j < my - 1 && k > 0 && k < mz - 1)) continue;
      rct.x[g.at(0, j, k)] = rct.x[g.at(mx - 2, j, k)];
      rct.x[g.at(mx - 1, j, k)] = rct.x[g.at(1, j, k)];
    } else {
      if (!g.zper) continue;
      long r = t - nyz;
      int i = (int)(r % mx), j = (int)(r / mx);
      if (!(i > 0 && i < mx - 1 && j > 0 && j < my - 1)) continue;
      rct.z[g.at(i, j, 0)] = rct.z[g.at(i, j, mz - 2)];
      rct.z[g.at(i, j, mz - 1)] = rct.z[g.at(i, j, 1)];
    }
  }
}
__global__ void k_rct_periodic(G g, V3 rct) { rct_periodic_body(g, rct, GS_SELF); }


// ------------------------------------------------------------------
// FormFunction1 face right-hand side (rhs.c:2799-3306)
// ------------------------------------------------------------------
__device__ inline int face_mode(const uint32_t *fl, long t, int d, int tt) {
  int q = (d == 0) ? (tt == 1 ? 0 : 1) : (d == 1) ? (tt == 0 ? 0 : 1) : (tt == 0 ? 0 : 1);
  return (fl[t] >> (2 + 5 * d + 2 * q)) & 3u;
}

// tangential derivative of p along tt on face (i,j,k) of axis d by mode
__device__ inline double ptan(const G &g, const double *p, int d, int tt, long t, int mode) {
  const long so[3] = {1, g.sj, g.sk};
  const long o = so[d], f = so[tt];
  if (mode == XM_LOWER) return (p[t] + p[t + o] - p[t - f] - p[t + o - f]) * 0.5;
  if (mode == XM_UPPER) return (p[t + f] + p[t + o + f] - p[t] - p[t + o]) * 0.5;
  return (p[t + f] + p[t + o + f] - p[t - f] - p[t + o - f]) * 0.25;   // CENTRAL
}

struct FFArgs {
  G g;
  const double *p, *nv;
  const uint32_t *fl;
  C3 rct;
  const double *gc[9];   // face coefficients (xi_d . xi_t) / J of the fine level, g[3d + t]
};

// rhs = avg(rct) - (dpdc g_d0 + dpde g_d1 + dpdz g_d2) on face t of axis d;
// g_dt = (xi_d . xi_t) * aj_face are the coefficients the Poisson operator
// uses (rhs.c:2869-2878 multiplies the same dot products by iaj after the
// sum; the two orders agree to rounding)
__device__ inline double face_rhs_value(const FFArgs &A, int d, long t, double dpdc, double dpde, double dpdz) {
  const double *rc = d == 0 ? A.rct.x : d == 1 ? A.rct.y : A.rct.z;
  const long so[3] = {1, A.g.sj, A.g.sk};
  double r = 0.5 * (rc[t] + rc[t + so[d]]);
  r -= dpdc * A.gc[3 * d][t] + dpde * A.gc[3 * d + 1][t] + dpdz * A.gc[3 * d + 2][t];
  return r;
}

// face_rhs_value with the two rct values of the face already loaded
__device__ inline double face_rhs_v(const FFArgs &A, int d, long t, double rlo, double rhi, double dpdc, double dpde,
                                    double dpdz) {
  double r = 0.5 * (rlo + rhi);
  r -= dpdc * A.gc[3 * d][t] + dpde * A.gc[3 * d + 1][t] + dpdz * A.gc[3 * d + 2][t];
  return r;
}

// the z-face (dpde, dpdz) state a cell leaves behind in loop order, used by
// the next cell's DROP carry
__device__ inline void zface_state(const FFArgs &A, long t, double *dpde, double *dpdz) {
  const G &g = A.g;
  int me = face_mode(A.fl, t, 2, 1);
  if (me == XM_DROP) {   // z-face dpde DROP keeps the y-face value p[j+1]-p[j]
    *dpde = A.p[t + g.sj] - A.p[t];
  } else {
    *dpde = ptan(g, A.p, 2, 1, t, me);
  }
  *dpdz = A.p[t + g.sk] - A.p[t];
}

__device__ inline long prev_cell(const G &g, int i, int j, int k) {
  if (i > 1) return g.at(i - 1, j, k);
  if (j > 1) return g.at(g.mx - 2, j - 1, k);
  if (k > 1) return g.at(g.mx - 2, g.my - 2, k - 1);
  return -1;
}

// ptan from loaded values: p0 = p[t], po = p[t+o], pf and pof = p[t+f],
// p[t+o+f], pmf and pomf = p[t-f], p[t+o-f]; the same terms in the same
// order as ptan, so the same bits
__device__ inline double ptan_v(int mode, double p0, double po, double pf, double pof, double pmf, double pomf) {
  const bool lo = mode == XM_LOWER, up = mode == XM_UPPER;
  const double a = lo ? p0 : pf, b = lo ? po : pof;
  const double c = up ? p0 : pmf, d = up ? po : pomf;
  return (a + b - c - d) * (lo || up ? 0.5 : 0.25);
}

__device__ __forceinline__ void ff1_faces_body(FFArgs A, V3 rhs, long n, int bx, int by, int bz) {
  const G &g = A.g;
  (void)n;
  // what the cell below (k-1) loaded that this cell reads again: its plane
  // k+1 points (this cell's p, pe, pn, pw, ps), its own p, pe, pn (this
  // cell's k-1 points), rct.z and nvert at k+1; valid after an interior cell
  bool have = false;
  double c_pt = 0., c_pet = 0., c_pnt = 0., c_ptw = 0., c_pts = 0., c_p0 = 0., c_pe = 0., c_pn = 0.;
  double c_rz = 0., c_nv = 0.;
  CV_MARCH_B(g, bx, by, bz) {
    const long t = g.at(i, j, k);
    if (!(i > 0 && i < g.mx - 1 && j > 0 && j < g.my - 1 && k > 0 && k < g.mz - 1)) { have = false; continue; }
    const double *p = A.p;
    const long sj = g.sj, sk = g.sk;
    const uint32_t w = A.fl[t];
    // mode of the tangential derivative q (0, 1: the two other axes in
    // order) on the face of axis d, as face_mode reads it
    auto fm = [w](int d, int q) { return (int)((w >> (2 + 5 * d + 2 * q)) & 3u); };
    // the 16 points of p the three faces' rules read (all inside the grid
    // for an interior cell), loaded before any mode test so that their
    // loads overlap; 8 of them come from the cell below when it was interior
    double p0, pe, pn, pw, ps, pb, pnb, peb, rz0, n0;
    if (have) {
      p0 = c_pt; pe = c_pet; pn = c_pnt; pw = c_ptw; ps = c_pts;
      pb = c_p0; pnb = c_pn; peb = c_pe;
      rz0 = c_rz; n0 = c_nv;
    } else {
      p0 = p[t]; pe = p[t + 1]; pn = p[t + sj]; pw = p[t - 1]; ps = p[t - sj];
      pb = p[t - sk]; pnb = p[t + sj - sk]; peb = p[t + 1 - sk];
      rz0 = A.rct.z[t]; n0 = A.nv[t];
    }
    const double pt = p[t + sk], pnt = p[t + sj + sk], pet = p[t + 1 + sk], ptw = p[t + sk - 1], pts = p[t + sk - sj];
    const double pen = p[t + 1 + sj], pes = p[t + 1 - sj], pnw = p[t + sj - 1];
    const double rz1 = A.rct.z[t + sk], nvt = A.nv[t + sk];
    // carried values from the previous cell in loop order
    double cde = 0.0, cdz = 0.0;
    const int mxe = fm(0, 0), mxz = fm(0, 1);
    if (mxe == XM_DROP || mxz == XM_DROP) {
      long pc = prev_cell(g, i, j, k);
      if (pc >= 0) zface_state(A, pc, &cde, &cdz);
    }
    // x face: dpdc set, dpde rule, dpdz rule
    double xdc = pe - p0;
    double xde = mxe == XM_DROP ? cde : ptan_v(mxe, p0, pe, pn, pen, ps, pes);
    double xdz = mxz == XM_DROP ? cdz : ptan_v(mxz, p0, pe, pt, pet, pb, peb);
    double rx = face_rhs_value(A, 0, t, xdc, xde, xdz);
    // y face: dpdc rule, dpde set, dpdz rule
    const int myc = fm(1, 0), myz = fm(1, 1);
    double ydc = myc == XM_DROP ? xdc : ptan_v(myc, p0, pn, pe, pen, pw, pnw);
    double yde = pn - p0;
    double ydz = myz == XM_DROP ? xdz : ptan_v(myz, p0, pn, pt, pnt, pb, pnb);
    double ry = face_rhs_value(A, 1, t, ydc, yde, ydz);
    // z face: dpdc rule, dpde rule, dpdz set
    const int mzc = fm(2, 0), mze = fm(2, 1);
    double zdc = mzc == XM_DROP ? ydc : ptan_v(mzc, p0, pt, pe, pet, pw, ptw);
    double zde = mze == XM_DROP ? yde : ptan_v(mze, p0, pt, pn, pnt, ps, pts);
    double zdz = pt - p0;
    double rz = face_rhs_v(A, 2, t, rz0, rz1, zdc, zde, zdz);
    // masks (rhs.c:3280-3306)
    const double *nv = A.nv;
    if (n0 > 0.1) { rx = 0; ry = 0; rz = 0; }
    if (nv[t + 1] > 0.1) rx = 0;
    if (nv[t + sj] > 0.1) ry = 0;
    if (nvt > 0.1) rz = 0;
    rhs.x[t] = rx; rhs.y[t] = ry; rhs.z[t] = rz;
    have = true;
    c_pt = pt; c_pet = pet; c_pnt = pnt; c_ptw = ptw; c_pts = pts;
    c_p0 = p0; c_pe = pe; c_pn = pn;
    c_rz = rz1; c_nv = nvt;
  }
}
__global__ void __launch_bounds__(BS) k_ff1_faces(FFArgs A, V3 rhs, long n) { ff1_faces_body(A, rhs, n, blockIdx.x, blockIdx.y, blockIdx.z); }


// the loop state after the main FormFunction1 sweep: the z-face values of
// the last interior cell, per variable (dpdc, dpde, dpdz)
__device__ inline void main_end_state(const FFArgs &A, double *dc, double *de, double *dz) {
  const G &g = A.g;
  long t = g.at(g.mx - 2, g.my - 2, g.mz - 2);
  const double *p = A.p;
  double pe = 0.0, pz = 0.0;
  int mxe = face_mode(A.fl, t, 0, 1), mxz = face_mode(A.fl, t, 0, 2);
  if (mxe == XM_DROP || mxz == XM_DROP) {
    long pc = prev_cell(g, g.mx - 2, g.my - 2, g.mz - 2);
    if (pc >= 0) zface_state(A, pc, &pe, &pz);
  }
  double xdc = p[t + 1] - p[t];
  int myc = face_mode(A.fl, t, 1, 0), myz = face_mode(A.fl, t, 1, 2);
  double ydc = myc == XM_DROP ? xdc : ptan(g, p, 1, 0, t, myc);
  double yde = p[t + g.sj] - p[t];
  (void)myz;
  int mzc = face_mode(A.fl, t, 2, 0), mze = face_mode(A.fl, t, 2, 1);
  *dc = mzc == XM_DROP ? ydc : ptan(g, p, 2, 0, t, mzc);
  *de = mze == XM_DROP ? yde : ptan(g, p, 2, 1, t, mze);
  *dz = p[t + g.sk] - p[t];
  (void)pe;
}

// periodic boundary faces (rhs.c:3032-3278): x faces at i=0 (k outer, j
// inner), then z faces at k=0 (j outer, i inner).  A DROP keeps the value of
// the previous face in that order, found by walking back.
__device__ __forceinline__ void ff1_periodic_body(FFArgs A, V3 rhs, GS gs) {
  const G &g = A.g;
  const int ny = g.my - 2, nxk = g.mx - 2, nzk = g.mz - 2;
  const long nxf = A.g.xper ? (long)ny * nzk : 0, nzf = A.g.zper ? (long)ny * nxk : 0;
  for (long t = gs.t0; t < nxf + nzf; t += gs.step) {
    const double *p = A.p;
    if (t < nxf) {
      int j = 1 + (int)(t % ny), k = 1 + (int)(t / ny);
      long f = g.at(0, j, k);
      double dc = p[f + 1] - p[f];
      double v[2];
      for (int q = 0; q < 2; q++) {        // q=0: dpde (tt=1), q=1: dpdz (tt=2)
        int tt = q == 0 ? 1 : 2;
        int mode = face_mode(A.fl, f, 0, tt);
        if (mode != XM_DROP) { v[q] = ptan(g, p, 0, tt, f, mode); continue; }
        // walk back along the periodic x-face sequence
        long s = t - 1;
        bool found = false;
        while (s >= 0) {
          int jj = 1 + (int)(s % ny), kk = 1 + (int)(s / ny);
          long ff = g.at(0, jj, kk);
          int mm = face_mode(A.fl, ff, 0, tt);
          if (mm != XM_DROP) { v[q] = ptan(g, p, 0, tt, ff, mm); found = true; break; }
          s--;
        }
        if (!found) {
          double a, b, c2;
          main_end_state(A, &a, &b, &c2);
          v[q] = q == 0 ? b : c2;
        }
      }
      rhs.x[f] = face_rhs_value(A, 0, f, dc, v[0], v[1]);
    } else {
      long r = t - nxf;
      int i = 1 + (int)(r % nxk), j = 1 + (int)(r / nxk);
      long f = g.at(i, j, 0);
      double dz = p[f + g.sk] - p[f];
      double v[2];
      for (int q = 0; q < 2; q++) {        // q=0: dpdc (tt=0), q=1: dpde (tt=1)
        int tt = q;
        int mode = face_mode(A.fl, f, 2, tt);
        if (mode != XM_DROP) { v[q] = ptan(g, p, 2, tt, f, mode); continue; }
        long s = r - 1;
        bool found = false;
        while (s >= 0) {
          int ii = 1 + (int)(s % nxk), jj = 1 + (int)(s / nxk);
          long ff = g.at(ii, jj, 0);
          int mm = face_mode(A.fl, ff, 2, tt);
          if (mm != XM_DROP) { v[q] = ptan(g, p, 2, tt, ff, mm); found = true; break; }
          s--;
        }
        if (!found) {
          if (A.g.xper) {   // last periodic x face: dpdc = p[1]-p[0], dpde its rule value
            long ff = g.at(0, g.my - 2, g.mz - 2);
            if (q == 0) v[q] = p[ff + 1] - p[ff];
            else {
              int mm = face_mode(A.fl, ff, 0, 1);
              double a, b, c2;
              main_end_state(A, &a, &b, &c2);
              v[q] = mm != XM_DROP ? ptan(g, p, 0, 1, ff, mm) : b;
            }
          } else {
            double a, b, c2;
            main_end_state(A, &a, &b, &c2);
            v[q] = q == 0 ? a : b;
          }
        }
      }
      rhs.z[f] = face_rhs_value(A, 2, f, v[0], v[1], dz);
    }
  }
}
__global__ void k_ff1_periodic(FFArgs A, V3 rhs) { ff1_periodic_body(A, rhs, GS_SELF); }


__device__ __forceinline__ void rk_update_body(double a, C3 uo, C3 r, V3 u, long n, GS gs) {
  for (long t = gs.t0; t < n; t += gs.step) {
    u.x[t] = a * r.x[t] + uo.x[t];
    u.y[t] = a * r.y[t] + uo.y[t];
    u.z[t] = a * r.z[t] + uo.z[t];
  }
}
__global__ void k_rk_update(double a, C3 uo, C3 r, V3 u, long n) { rk_update_body(a, uo, r, u, n, GS_SELF); }



// CalcRHS (implicitsolver.c:6449-6741) after FormFunction1: Rhs += (-1/dt)
// du with du = 1.5 u - 2 u_o + 0.5 u_rm1 (BDF2) or u - u_o (first step),
// at every storage point, then the zero set of the reference's interior and
// boundary-plane masks (all of them only zero entries, so they commute)
__global__ void __launch_bounds__(BS) k_calcrhs(G g, int bdf2, double idt, C3 u, C3 uo, C3 um, const double *nv,
                                                V3 r, int yper) {
  const int mx = g.mx, my = g.my, mz = g.mz, xper = g.xper, zper = g.zper;
  CV_MARCH(g) {
    const long t = g.at(i, j, k);
    double d0, d1, d2;
    if (bdf2) {
      d0 = u.x[t] * 1.5; d0 = d0 + (-2.) * uo.x[t]; d0 = d0 + 0.5 * um.x[t];
      d1 = u.y[t] * 1.5; d1 = d1 + (-2.) * uo.y[t]; d1 = d1 + 0.5 * um.y[t];
      d2 = u.z[t] * 1.5; d2 = d2 + (-2.) * uo.z[t]; d2 = d2 + 0.5 * um.z[t];
    } else {
      d0 = -1. * uo.x[t] + u.x[t];
      d1 = -1. * uo.y[t] + u.y[t];
      d2 = -1. * uo.z[t] + u.z[t];
    }
    double rx = r.x[t] + idt * d0, ry = r.y[t] + idt * d1, rz = r.z[t] + idt * d2;
    bool zx = false, zy = false, zz = false;
    const long sj = g.sj, sk = g.sk;
    auto nvx = [&]() { return nv[t] + nv[t + 1] > 0.1; };
    auto nvy = [&]() { return nv[t] + nv[t + sj] > 0.1; };
    auto nvz = [&]() { return nv[t] + nv[t + sk] > 0.1; };
    // interior cells (6524-6538)
    if (i >= 1 && i <= mx - 2 && j >= 1 && j <= my - 2 && k >= 1 && k <= mz - 2) {
      if (nvx() || (i > mx - 3 && !xper)) zx = true;
      if (nvy() || (j > my - 3 && !yper)) zy = true;
      if (nvz() || (k > mz - 3 && !zper)) zz = true;
    }
    // i planes (6541-6583): j, k in [0, m-2]
    if (j <= my - 2 && k <= mz - 2) {
      if (i == 0) {
        if (!xper) zx = zy = zz = true;
        else { if (nvx()) zx = true; if (nvy()) zy = true; if (nvz()) zz = true; }
      }
      if (i == mx - 2 && !xper) zx = true;
      if (i == mx - 1) zx = zy = zz = true;
    }
    // j planes (6587-6636): i, k in [0, m-2]
    if (i <= mx - 2 && k <= mz - 2) {
      if (j == 0) {
        if (!yper) zx = zy = zz = true;
        else { if (nvx()) zx = true; if (nvy()) zy = true; if (nvz()) zz = true; }
      }
      if (j == my - 2 && !yper) zy = true;
      if (j == my - 1) zx = zy = zz = true;
    }
    // k planes (6639-6690): i, j in [0, m-2]
    if (i <= mx - 2 && j <= my - 2) {
      if (k == 0) {
        if (!zper) zx = zy = zz = true;
        else { if (nvx()) zx = true; if (nvy()) zy = true; if (nvz()) zz = true; }
      }
      if (k == mz - 2 && !zper) zz = true;
      if (k == mz - 1) zx = zy = zz = true;
    }
    r.x[t] = zx ? 0. : rx;
    r.y[t] = zy ? 0. : ry;
    r.z[t] = zz ? 0. : rz;
  }
}

// max |a - b| (diff) or max |a| over three planar arrays; part[block]
__global__ void __launch_bounds__(BS) k_absmax3(C3 a, C3 b, int diff, long n, double *part) {
  __shared__ double sh[BS / 64];
  double m = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    double v0 = a.x[t], v1 = a.y[t], v2 = a.z[t];
    if (diff) { v0 = b.x[t] - v0; v1 = b.y[t] - v1; v2 = b.z[t] - v2; }
    m = cv_nmax(m, cv_nmax(fabs(v0), cv_nmax(fabs(v1), fabs(v2))));
  }
  m = cv_block_max(m, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void k_final_max(const double *part, int nb, double *out) {
  __shared__ double sh[BS / 64];
  double m = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) m = cv_nmax(m, part[q]);
  m = cv_block_max(m, sh);
  if (threadIdx.x == 0) out[0] = m;
}
}  // namespace

// ====================================================================
// host side
// ====================================================================
static G gview(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  G g;
  g.mx = F.d.mx; g.my = F.d.my; g.mz = F.d.mz; g.xper = F.d.xper; g.zper = F.d.zper;
  g.sj = g.mx; g.sk = (long)g.mx * g.my;
  return g;
}
static V3 v3(double *const a[3]) { return V3{a[0], a[1], a[2]}; }
static C3 c3(double *const a[3]) { return C3{a[0], a[1], a[2]}; }

int cv_contra2cart_dev(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  CvMetrics &M = c->m;
  hipLaunchKernelGGL(k_contra2cart, march_grid(F.d.mx, F.d.my, F.d.mz), MARCH_BLOCK, 0, c->stream, gview(c), c3(c->ucont), c3(M.csi),
                     c3(M.eta), c3(M.zet), (const double *)F.nvert, v3(c->ucat), F.d.n);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// blocks of a plane kernel over the (mx-2) x (my-2) interior of a k plane
static unsigned plane_nb(const G &g) {
  const long m2 = (long)(g.mx - 2) * (g.my - 2);
  return (unsigned)std::min(256L, std::max(1L, (m2 + 4 * BS - 1) / (4 * BS)));
}

// InflowFlux (bcs.c:99-246): inlet fluxes and Ubcs of a bctype-5 k- face,
// FluxInSum to dsc (device); nothing without an inlet, as FluxInSum is
// only written by the inlet branch
int cv_inflow_dev(cv_ctx *c) {
  if (!c->inlet) return CV_OK;
  CvLevel &F = c->lev[c->nlev - 1];
  const G g = gview(c);
  const unsigned nb = plane_nb(g);
  const Inlet in{c->inletprofile, c->cm_c[0], c->cm_c[1], c->cm_c[2]};
  hipLaunchKernelGGL(k_inflow, dim3(nb), dim3(BS), 0, c->stream, g, in, (const double *)F.coor, F.d.n, c3(c->m.zet),
                     (const double *)F.nvert, c->ucont[2], c->ubf, c->part);
  hipLaunchKernelGGL(cv_k_final_sum, dim3(1), dim3(BS), 0, c->stream, (const double *)c->part, (int)nb,
                     c->dsc + CV_DSC_FLUXIN, 1);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// FormBCS's outlet section (bcs.c:4914-5010) from the current Ucat; with
// stat the corrected outflow FluxOutSum to dsc
static void outlet_fix(cv_ctx *c, int stat) {
  CvLevel &F = c->lev[c->nlev - 1];
  const G g = gview(c);
  const unsigned nb = plane_nb(g);
  const double *nv = F.nvert;
  hipLaunchKernelGGL(k_outlet_sum, dim3(nb), dim3(BS), 0, c->stream, g, c3(c->ucat), c3(c->m.zet), nv, c->part);
  hipLaunchKernelGGL(cv_k_final_sum, dim3(2), dim3(BS), 0, c->stream, (const double *)c->part, (int)nb,
                     c->dsc + CV_DSC_OUTSUM, 2);
  hipLaunchKernelGGL(k_outlet_fix, dim3(nb), dim3(BS), 0, c->stream, g, c3(c->ucat), c3(c->m.zet), nv, c->ucont[2],
                     c->ubf, (const double *)c->dsc, stat, c->part + 2 * nb);
  if (stat)
    hipLaunchKernelGGL(cv_k_final_sum, dim3(1), dim3(BS), 0, c->stream, (const double *)(c->part + 2 * nb), (int)nb,
                       c->dsc + CV_DSC_FLUXOUT, 1);
}

// FormBCS's far-field section (bcs.c:4500-4724): Ubcs, fluxes and sums,
// then the flux balance
static void far_fix(cv_ctx *c) {
  const G g = gview(c);
  CvMetrics &M = c->m;
  const long np = face_points(g.mx, g.my, g.mz);
  const unsigned nb = (unsigned)std::min(256L, std::max(1L, (np + 4 * BS - 1) / (4 * BS)));
  hipLaunchKernelGGL(k_far_sum, dim3(nb), dim3(BS), 0, c->stream, g, c->far, c3(c->ucat), c3(M.csi), c3(M.eta),
                     c3(M.zet), v3(c->ucont), c->ubf, c->part);
  hipLaunchKernelGGL(cv_k_final_sum, dim3(5), dim3(BS), 0, c->stream, (const double *)c->part, (int)nb,
                     c->dsc + CV_DSC_FAR, 5);
  const int has_out = (c->far & ((1 << 1) | (1 << 3) | (1 << 5))) != 0;
  hipLaunchKernelGGL(k_far_fix, dim3(nb), dim3(BS), 0, c->stream, g, c->far, has_out, c3(M.csi), c3(M.eta), c3(M.zet),
                     v3(c->ucont), c->ubf, (const double *)c->dsc);
}

// FormBCS's slip section (formbcs) or GhostNodeVelocity's
static void slip_update(cv_ctx *c, int formbcs) {
  const G g = gview(c);
  CvMetrics &M = c->m;
  hipLaunchKernelGGL(k_slip, dim3(nblk(face_points(g.mx, g.my, g.mz))), dim3(BS), 0, c->stream, g, c->slip, formbcs,
                     c3(c->ucat), c3(M.csi), c3(M.eta), c3(M.zet), v3(c->ucont), c->ubf);
}

// formbcs=1: the three passes of FormBCS, each far field -> slip -> outlet
// -> Contra2Cart -> ghosts.  Contra2Cart runs over the whole grid in the
// first pass; later passes repeat it only where the pass changed Ucont:
// the whole grid with far-field faces, the k = mz-2 plane with an outlet
// alone (the slip faces' zero fluxes are already in place).
// Only a far-field, slip or outlet face changes Ucont or Ubcs inside a
// pass; without one, the first pass leaves a fixed point of the pass (wall
// ghosts and the periodic plane copies take interior values, the edge
// averages those face ghosts, and each corner's last write edges of the same
// pass), so one pass gives the three passes' bits.
// formbcs=0: GhostNodeVelocity (slip Ubcs, then the ghosts)
// cart=false: the interior Ucat of the first pass is already in place
// (k_rk_c2c formed it with the RK stage update)
int cv_ghost_update(cv_ctx *c, int formbcs, bool cart) {
  CvLevel &F = c->lev[c->nlev - 1];
  CvMetrics &M = c->m;
  G g = gview(c);
  const int bxper = F.d.xper, bzper = F.d.zper;
  const long planes = (long)g.my * g.mz + (long)g.mx * g.mz + (long)g.mx * g.my;
  const int passes = formbcs && (c->far || c->slip || c->outlet) ? 3 : 1;
  const bool outl = formbcs && c->outlet, far = formbcs && c->far;
  double *p = formbcs ? c->p : nullptr, *nv = formbcs ? F.nvert : nullptr;
  for (int ps = 0; ps < passes; ps++) {
    if (far) far_fix(c);
    if (c->slip) slip_update(c, formbcs);
    if (outl) outlet_fix(c, ps == passes - 1);
    if (formbcs && ((ps == 0 && cart) || far)) {
      int rc = cv_contra2cart_dev(c);
      if (rc) return rc;
    } else if (outl) {
      hipLaunchKernelGGL(k_contra2cart_plane, dim3(plane_nb(g)), dim3(BS), 0, c->stream, g, c3(c->ucont), c3(M.csi),
                         c3(M.eta), c3(M.zet), (const double *)F.nvert, v3(c->ucat), g.mz - 2);
    }
    hipLaunchKernelGGL(k_wall_ghost, dim3(nblk(planes)), dim3(BS), 0, c->stream, g, bxper, bzper, formbcs ? 0 : 1, v3(c->ucat),
                       (const double *)c->ubf);
    if (bxper || bzper)
      hipLaunchKernelGGL(k_periodic_planes, dim3(nblk((long)g.my * g.mz + (long)g.mx * g.my)), dim3(BS), 0, c->stream, g,
                         v3(c->ucat), p, nv);
    // one block runs the twelve ordered sub-sweeps; 1024 threads cover an
    // edge line of up to 1024 points in one pass per sub-sweep
    hipLaunchKernelGGL(k_edge_average, dim3(1), dim3(1024), 0, c->stream, g, v3(c->ucat), p);
    if (bxper)
      hipLaunchKernelGGL(k_periodic_full, dim3(nblk(2L * g.mz + 2L * (g.my - 2))), dim3(BS), 0, c->stream, g, 0, formbcs ? 1 : 0,
                         v3(c->ucat), p, nv);
    if (bzper)
      hipLaunchKernelGGL(k_periodic_full, dim3(nblk(2L * g.my + 2L * (g.mx - 2))), dim3(BS), 0, c->stream, g, 2, 0, v3(c->ucat), p, nv);
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// the boundary update before and after every RK stage: InflowFlux,
// OutflowFlux (a diagnostic whose sum FormBCS overwrites before any use, so
// not launched here) and FormBCS (implicitsolver.c:43-46, 79-84)
int cv_bc_update_dev(cv_ctx *c) {
  int rc = cv_inflow_dev(c);
  return rc ? rc : cv_ghost_update(c, 1);
}

// ---- immersed-boundary nodes ----
void cv_ib_clear(cv_ctx *c) {
  void *bufs[] = {c->ib_node, c->ib_w, c->ib_new, c->ib_face, c->ib_cell};
  for (void *b : bufs)
    if (b) cv_free(c, b);
  c->ib_node = nullptr; c->ib_w = nullptr; c->ib_new = nullptr; c->ib_face = nullptr; c->ib_cell = nullptr;
  c->ib_n = 0; c->ib_ncell = 0;
  c->ib_nface[0] = c->ib_nface[1] = c->ib_nface[2] = 0;
  c->ib_gen++;
}

// six sweeps of node reconstruction -> face fluxes -> Contra2Cart.  p is
// rebuilt in the first sweep only: every sweep of the reference computes
// it from the same lP, the P at entry.
int cv_ib_dev(cv_ctx *c) { return cv_ib_dev_on(c, c->stream); }

int cv_ib_dev_on(cv_ctx *c, hipStream_t stream) {
  if (c->ib_n <= 0) return CV_OK;
  CvLevel &F = c->lev[c->nlev - 1];
  CvMetrics &M = c->m;
  const G g = gview(c);
  const int n = c->ib_n, nf = c->ib_nface[0] + c->ib_nface[1] + c->ib_nface[2];
  for (int sweep = 0; sweep < 6; sweep++) {
    const int with_p = sweep == 0;
    hipLaunchKernelGGL(k_ib_nodes, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, n, (const int *)c->ib_node,
                       (const double *)c->ib_w, c->ib_mode, c->dt, c3(c->ucat), (const double *)c->p, with_p, c->ib_new);
    hipLaunchKernelGGL(k_ib_scatter, dim3((n + BS - 1) / BS), dim3(BS), 0, stream, n, (const int *)c->ib_node,
                       (const double *)c->ib_new, v3(c->ucat), c->p, with_p);
    if (nf > 0)
      hipLaunchKernelGGL(k_ib_faces, dim3((nf + BS - 1) / BS), dim3(BS), 0, stream, g, c->ib_nface[0], c->ib_nface[1],
                         c->ib_nface[2], (const int *)c->ib_face, c3(c->ucat), c3(M.csi), c3(M.eta), c3(M.zet),
                         v3(c->ucont));
    if (c->ib_ncell > 0)
      hipLaunchKernelGGL(k_ib_c2c, dim3((c->ib_ncell + BS - 1) / BS), dim3(BS), 0, stream, g, c->ib_ncell,
                         (const int *)c->ib_cell, c3(c->ucont), c3(M.csi), c3(M.eta), c3(M.zet), (const double *)F.nvert,
                         v3(c->ucat));
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

extern "C" int cv_set_ib_nodes(cv_ctx *c, const cv_ib_node *nodes, int n, int mode) {
  if (!c || n < 0 || (n > 0 && !nodes) || (mode != CV_IB_NOSLIP && mode != CV_IB_FREESLIP)) return CV_ERR_ARG;
  if (!c->geometry_set || !c->nvert_set) { cv_set_error("cv_set_ib_nodes: geometry and nvert must be set first"); return CV_ERR_STATE; }
  CV_CHECK(hipSetDevice(c->device));
  CV_CHECK(hipStreamSynchronize(c->stream));
  cv_ib_clear(c);
  if (n == 0) return CV_OK;
  CvLevel &F = c->lev[c->nlev - 1];
  const int mx = F.d.mx, my = F.d.my, mz = F.d.mz;
  const long N = F.d.n, sj = mx, sk = (long)mx * my;
  auto at = [&](int i, int j, int k) { return (long)i + mx * ((long)j + (long)my * k); };
  auto inside = [&](int i, int j, int k, int lo) { return i >= lo && i < mx - lo && j >= lo && j < my - lo && k >= lo && k < mz - lo; };
  std::vector<int> node(4 * (size_t)n);
  std::vector<double> w(12 * (size_t)n);
  for (int q = 0; q < n; q++) {
    const cv_ib_node &b = nodes[q];
    if (!inside(b.i, b.j, b.k, 1) || !inside(b.i1, b.j1, b.k1, 0) || !inside(b.i2, b.j2, b.k2, 0) ||
        !inside(b.i3, b.j3, b.k3, 0)) {
      cv_set_error("cv_set_ib_nodes: node %d or its interception cells lie outside the grid", q);
      return CV_ERR_ARG;
    }
    if (!(b.d_s >= 0.0) || !(b.d_s + b.d_i > 0.0)) {
      cv_set_error("cv_set_ib_nodes: node %d needs d_s >= 0 and d_s + d_i > 0", q);
      return CV_ERR_ARG;
    }
    node[4 * q] = (int)at(b.i, b.j, b.k);
    node[4 * q + 1] = (int)at(b.i1, b.j1, b.k1);
    node[4 * q + 2] = (int)at(b.i2, b.j2, b.k2);
    node[4 * q + 3] = (int)at(b.i3, b.j3, b.k3);
    const double v[12] = {b.cr1, b.cr2, b.cr3, b.d_s, b.d_i, b.ua[0], b.ua[1], b.ua[2], b.dun, b.nf[0], b.nf[1], b.nf[2]};
    for (int e = 0; e < 12; e++) w[12 * (size_t)q + e] = v[e];
  }
  // faces touching an IB node ((int)(nvert + 0.5) == 1, ibm.c:3166-3224) and
  // the cells they bound, from the device nvert
  std::vector<double> nv(N);
  CV_CHECK(hipMemcpyAsync(nv.data(), F.nvert, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  auto ib = [&](long t) { return (int)(nv[t] + 0.5) == 1; };
  std::vector<int> fl[3];
  std::vector<unsigned char> mark(N, 0);
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        const long t = at(i, j, k);
        const bool b0 = ib(t);
        if (i < mx - 2 && (b0 || ib(t + 1))) { fl[0].push_back((int)t); mark[t] = mark[t + 1] = 1; }
        if (j < my - 2 && (b0 || ib(t + sj))) { fl[1].push_back((int)t); mark[t] = mark[t + sj] = 1; }
        if (k < mz - 2 && (b0 || ib(t + sk))) { fl[2].push_back((int)t); mark[t] = mark[t + sk] = 1; }
      }
  std::vector<int> faces, cells;
  for (int d = 0; d < 3; d++) faces.insert(faces.end(), fl[d].begin(), fl[d].end());
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++)
        if (mark[at(i, j, k)]) cells.push_back((int)at(i, j, k));
  void *p;
  int rc;
  if ((rc = cv_alloc(c, &p, node.size() * sizeof(int)))) return rc;
  c->ib_node = (int *)p;
  if ((rc = cv_alloc(c, &p, w.size() * sizeof(double)))) return rc;
  c->ib_w = (double *)p;
  if ((rc = cv_alloc(c, &p, 4 * (size_t)n * sizeof(double)))) return rc;
  c->ib_new = (double *)p;
  if ((rc = cv_alloc(c, &p, std::max<size_t>(1, faces.size()) * sizeof(int)))) return rc;
  c->ib_face = (int *)p;
  if ((rc = cv_alloc(c, &p, std::max<size_t>(1, cells.size()) * sizeof(int)))) return rc;
  c->ib_cell = (int *)p;
  CV_CHECK(hipMemcpy(c->ib_node, node.data(), node.size() * sizeof(int), hipMemcpyHostToDevice));
  CV_CHECK(hipMemcpy(c->ib_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  if (!faces.empty()) CV_CHECK(hipMemcpy(c->ib_face, faces.data(), faces.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!cells.empty()) CV_CHECK(hipMemcpy(c->ib_cell, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
  for (int d = 0; d < 3; d++) c->ib_nface[d] = (int)fl[d].size();
  c->ib_ncell = (int)cells.size();
  c->ib_n = n;
  c->ib_mode = mode;
  return CV_OK;
}

// rows per face-flux block (4 or 8, CV_RC_ROWS; row 0 is halo); 8 is the
// faster at 512x256x256
static int rc_rows() {
  static int ty = 0;
  if (!ty) {
    const char *e = getenv("CV_RC_ROWS");
    ty = (e && atoi(e) == 4) ? 4 : 8;
  }
  return ty;
}

// k planes per face-flux block (CV_RC_KC)
static int rc_kchunk() {
  static int kc = 0;
  if (!kc) {
    const char *e = getenv("CV_RC_KC");
    kc = e ? std::max(1, atoi(e)) : RC_KC;
  }
  return kc;
}

// face-flux pass variant (CV_RC_KERNEL: 1 LDS-tiled, 0 global gathers),
// read at every call so one process can run both
// CV_RK_FUSED=0: the RK stage update, the wall flux reset and Contra2Cart
// as three passes also where k_rk_c2c applies (read per call)
static bool rk_fused() {
  const char *e = getenv("CV_RK_FUSED");
  return e ? atoi(e) != 0 : true;
}

static int rc_kernel() {
  const char *e = getenv("CV_RC_KERNEL");
  return e ? atoi(e) != 0 : 1;
}

static FFArgs ffargs(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  FFArgs A;
  A.g = gview(c);
  A.p = c->p; A.nv = F.nvert; A.fl = F.flags;
  A.rct = c3(c->rct);
  for (int q = 0; q < 9; q++) A.gc[q] = F.g[q];
  return A;
}

// FormFunction1: Contra2Cart is skipped (FormBCS already left Ucat of the
// same Ucont in place; the reference recomputes the identical values)
int cv_ff1_dev(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  CvMetrics &M = c->m;
  FFArgs A = ffargs(c);
  const double nu = 1. / c->ren;
  const C3 uc = c3(c->ucont), u = c3(c->ucat);
  const FM mi{c3(M.icsi), c3(M.ieta), c3(M.izet), M.iaj};
  const FM mj{c3(M.jcsi), c3(M.jeta), c3(M.jzet), M.jaj};
  const FM mk{c3(M.kcsi), c3(M.keta), c3(M.kzet), M.kaj};
  {
    // algorithmic bytes per point: 30 face metrics, 9 cell metrics + aj,
    // ucont, ucat, nvert, rct out
    CvProfScope ps(c, CV_PROF_RCT, c->nlev - 1, 0, (double)F.d.n * 8.0 * (30 + 10 + 3 + 3 + 1 + 3));
    const int ty = rc_rows(), tiled = rc_kernel();
    const int kc = rc_kchunk();
    const dim3 grid((unsigned)((F.d.mx + 62) / 63), (unsigned)((F.d.my - 1 + ty - 2) / (ty - 1)),
                    (unsigned)((F.d.mz + kc - 1) / kc));
    auto kern = tiled ? (ty == 8 ? k_rc_tiled<8> : k_rc_tiled<4>) : (ty == 8 ? k_rc_fused<8> : k_rc_fused<4>);
    hipLaunchKernelGGL(kern, grid, dim3(64, ty), 0, c->stream, A.g, uc, u, (const double *)F.nvert, mi, mj, mk, nu,
                       c3(M.csi), c3(M.eta), c3(M.zet), (const double *)M.aj, v3(c->rct), kc);
  }
  if (F.d.xper || F.d.zper)
    hipLaunchKernelGGL(k_rct_periodic, dim3(nblk((long)A.g.my * A.g.mz + (long)A.g.mx * A.g.my)), dim3(BS), 0, c->stream,
                       A.g, v3(c->rct));
  {
    // algorithmic bytes per point: p, rct, 9 face coefficients, nvert, Rhs out, the operator word
    CvProfScope ps(c, CV_PROF_FF1, c->nlev - 1, 0, (double)F.d.n * (8.0 * (1 + 3 + 9 + 1 + 3) + 4.0));
    hipLaunchKernelGGL(k_ff1_faces, march_grid(F.d.mx, F.d.my, F.d.mz), MARCH_BLOCK, 0, c->stream, A, v3(c->rhs), F.d.n);
  }
  if (F.d.xper || F.d.zper) {
    long nf = (F.d.xper ? (long)(A.g.my - 2) * (A.g.mz - 2) : 0) + (F.d.zper ? (long)(A.g.my - 2) * (A.g.mx - 2) : 0);
    hipLaunchKernelGGL(k_ff1_periodic, dim3(nblk(nf)), dim3(BS), 0, c->stream, A, v3(c->rhs));
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

int cv_rk_dev(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  G g = gview(c);
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  int rc = cv_bc_update_dev(c);
  if (rc || (rc = cv_ib_dev(c))) return rc;   // implicitsolver.c:47-51
  for (int q = 0; q < 3; q++)
    hipLaunchKernelGGL(k_fill, dim3(nblk(F.d.n)), dim3(BS), 0, c->stream, c->rhs[q], 0.0, F.d.n);
  const int wx = wall_mask(c->bctype[0], c->bctype[1]), wy = wall_mask(c->bctype[2], c->bctype[3]),
            wz = wall_mask(c->bctype[4], c->bctype[5]);
  const long planes = (long)g.my * g.mz + (long)g.mx * g.mz + (long)g.mx * g.my;
  // faces without an inlet, far-field, slip or outlet: FormBCS changes
  // only ghosts, so the stage update, the wall flux reset and Contra2Cart
  // run as one k-march and FormBCS makes its ghost pass alone
  const bool fused = !c->inlet && !c->far && !c->slip && !c->outlet && rk_fused();
  for (int s = 0; s < 4; s++) {
    if ((rc = cv_ff1_dev(c))) return rc;
    double a = alfa[s] * c->dt * c->st;
    if (fused) {
      CvMetrics &M = c->m;
      hipLaunchKernelGGL(k_rk_c2c, march_grid(F.d.mx, F.d.my, F.d.mz), MARCH_BLOCK, 0, c->stream, g, a, c3(c->ucont_o),
                         c3(c->rhs), v3(c->ucont), wx, wy, wz, c3(M.csi), c3(M.eta), c3(M.zet), (const double *)F.nvert,
                         v3(c->ucat));
      if ((rc = cv_ghost_update(c, 1, false))) return rc;
      continue;
    }
    hipLaunchKernelGGL(k_rk_update, dim3(nblk(F.d.n)), dim3(BS), 0, c->stream, a, c3(c->ucont_o), c3(c->rhs),
                       v3(c->ucont), F.d.n);
    hipLaunchKernelGGL(k_wall_flux_zero, dim3(nblk(planes)), dim3(BS), 0, c->stream, g, wx, wy, wz, v3(c->ucont));
    if ((rc = cv_bc_update_dev(c))) return rc;
  }
  if ((rc = cv_ib_dev(c))) return rc;         // after the stages, implicitsolver.c:87-90
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// CalcRHS: FormFunction1 into Rhs, then the time derivative and the masks
int cv_calcrhs_dev(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  G g = gview(c);
  int rc = cv_ff1_dev(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_calcrhs, march_grid(F.d.mx, F.d.my, F.d.mz), MARCH_BLOCK, 0, c->stream, g,
                     c->steps_done > 0 ? 1 : 0, -1. / c->dt, c3(c->ucont), c3(c->ucont_o), c3(c->ucont_rm1),
                     (const double *)F.nvert, v3(c->rhs), c->bctype[2] == 7 ? 1 : 0);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// max |a - b| (b non-null) or max |a| over the fine level, to the host
static int absmax3(cv_ctx *c, double *const a[3], double *const b[3], double *out) {
  const long n = c->lev[c->nlev - 1].d.n;
  const C3 ca = c3(a), cb = b ? c3(b) : ca;
  hipLaunchKernelGGL(k_absmax3, dim3(CV_RED_GRID), dim3(BS), 0, c->stream, ca, cb, b ? 1 : 0, n, c->part);
  hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BS), 0, c->stream, (const double *)c->part, CV_RED_GRID, c->dsc + 120);
  CV_CHECK(hipMemcpyAsync(c->red_host, c->dsc + 120, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  *out = c->red_host[0];
  return CV_OK;
}

static void copy3(cv_ctx *c, double *const dst[3], double *const src[3]) {
  const long n = c->lev[c->nlev - 1].d.n;
  for (int q = 0; q < 3; q++)
    hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(BS), 0, c->stream, dst[q], (const double *)src[q], n);
}

// ImpRK (implicitsolver.c:7094-7256), one block: pseudo-time RK stages on
// CalcRHS from the base pUcont, stop on ||pU - U||_inf (absolute or
// relative to the first pseudo iteration), halve lambda and reject the
// iteration when both ||Rhs||_inf and ||dU||_inf grew
int cv_imprk_dev(cv_ctx *c, cv_step_stats *st) {
  CvLevel &F = c->lev[c->nlev - 1];
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  int rc = cv_bc_update_dev(c);
  if (rc || (rc = cv_ib_dev(c))) return rc;   // implicitsolver.c:7128-7131
  for (int q = 0; q < 3; q++)
    hipLaunchKernelGGL(k_fill, dim3(nblk(F.d.n)), dim3(BS), 0, c->stream, c->rhs[q], 0.0, F.d.n);
  copy3(c, c->pucont, c->ucont);
  if ((rc = cv_calcrhs_dev(c))) return rc;
  double normF0, normF1 = 1000., normdU1 = 1000.;
  if ((rc = absmax3(c, c->rhs, nullptr, &normF0))) return rc;
  double lambda = c->mom.cfl;
  double normdU = 10., reldU = 1., normdU0 = 0., normdU_bk = 0., reldU_bk = 1., normF_bk = 0.;
  int pseudot = 0;
  while (((normdU > c->mom.imp_atol && reldU > c->mom.imp_rtol) || pseudot < 1) && pseudot < c->mom.imp_max_it) {
    pseudot++;
    for (int s = 0; s < 4; s++) {
      const double a = lambda * alfa[s] * c->dt * c->st;
      hipLaunchKernelGGL(k_rk_update, dim3(nblk(F.d.n)), dim3(BS), 0, c->stream, a, c3(c->pucont), c3(c->rhs),
                         v3(c->ucont), F.d.n);
      if ((rc = cv_bc_update_dev(c))) return rc;
      if ((rc = cv_calcrhs_dev(c))) return rc;
    }
    if ((rc = cv_ib_dev(c))) return rc;       // after the stages, implicitsolver.c:7170-7173
    normdU1 = normdU_bk;
    if ((rc = absmax3(c, c->ucont, c->pucont, &normdU_bk))) return rc;   // dU = pU - U
    if (pseudot == 1) { normdU0 = normdU_bk; reldU_bk = 1.; }
    else if (normdU0 > 1.e-10) reldU_bk = normdU_bk / normdU0;
    if (pseudot > 1) normF1 = normF_bk;
    if ((rc = absmax3(c, c->rhs, nullptr, &normF_bk))) return rc;
    if (pseudot == 1) normF0 = normF_bk;
    normdU = normdU_bk;
    reldU = reldU_bk;
    // an overflowing pseudo iteration (infinite norms) is rejected by the
    // lambda-halving test below like any growing one, as in the reference
    if (pseudot > 1 && normF_bk > normF1 && normdU_bk > normdU1 && lambda > 0.001) {
      lambda = 0.5 * lambda;
      pseudot--;
      copy3(c, c->ucont, c->pucont);
    } else {
      copy3(c, c->pucont, c->ucont);
      normF1 = normF_bk;
      normdU1 = normdU_bk;
    }
  }
  (void)normF0;
  if (st) { st->imp_its = pseudot; st->imp_lambda = lambda; st->imp_normdU = normdU; st->imp_reldU = reldU; }
  CV_CHECK(hipGetLastError());
  if (!std::isfinite(normdU)) {
    cv_set_error("ImpRK: pseudo iterations ended on a non-finite update");
    return CV_ERR_NONFINITE;
  }
  return CV_OK;
}

// ====================================================================
// C ABI
// ====================================================================
static int ready(cv_ctx *c) {
  if (!c) return CV_ERR_ARG;
  if (!c->geometry_set || !c->nvert_set) { cv_set_error("cv: geometry and nvert must be set first"); return CV_ERR_STATE; }
  CV_CHECK(hipSetDevice(c->device));
  return CV_OK;
}

extern "C" int cv_contra2cart(cv_ctx *c) {
  int rc = ready(c);
  if (rc) return rc;
  rc = cv_contra2cart_dev(c);
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cv_ibm_interpolation(cv_ctx *c) {
  int rc = ready(c);
  if (rc) return rc;
  rc = cv_ib_dev(c);
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cv_inflow_flux(cv_ctx *c, double *flux_in) {
  int rc = ready(c);
  if (rc) return rc;
  if ((rc = cv_inflow_dev(c))) return rc;
  CV_CHECK(hipMemcpyAsync(c->red_host + 110, c->dsc + CV_DSC_FLUXIN, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  if (flux_in) *flux_in = c->red_host[110];
  return CV_OK;
}

extern "C" int cv_outflow_flux(cv_ctx *c, double *flux_out) {
  int rc = ready(c);
  if (rc) return rc;
  double v = 0.0;
  if (c->outlet) {
    const G g = gview(c);
    const unsigned nb = plane_nb(g);
    hipLaunchKernelGGL(k_outflow_diag, dim3(nb), dim3(BS), 0, c->stream, g, (const double *)c->ucont[2], c->part);
    hipLaunchKernelGGL(cv_k_final_sum, dim3(1), dim3(BS), 0, c->stream, (const double *)c->part, (int)nb,
                       c->dsc + CV_DSC_OUTDIAG, 1);
    CV_CHECK(hipGetLastError());
    CV_CHECK(hipMemcpyAsync(c->red_host + 112, c->dsc + CV_DSC_OUTDIAG, sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    v = c->red_host[112];
  }
  if (flux_out) *flux_out = v;
  return CV_OK;
}

extern "C" int cv_form_bcs(cv_ctx *c) {
  int rc = ready(c);
  if (rc) return rc;
  rc = cv_ghost_update(c, 1);
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cv_form_function1(cv_ctx *c, double *rhs_out) {
  int rc = ready(c);
  if (rc) return rc;
  if ((rc = cv_contra2cart_dev(c))) return rc;
  if ((rc = cv_ff1_dev(c))) return rc;
  CV_CHECK(hipStreamSynchronize(c->stream));
  if (rhs_out) return cv_get_field(c, CV_F_RHS, -1, rhs_out);
  return CV_OK;
}

extern "C" int cv_rungekutta(cv_ctx *c) {
  int rc = ready(c);
  if (rc) return rc;
  rc = cv_rk_dev(c);
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

// the part of a fractional step before the pressure solve: the level
// rotation Ucont_rm1 <- Ucont_o <- Ucont (main.c:1131-1132, done at the step
// start) and the momentum scheme
int cv_step_pre_dev(cv_ctx *c, cv_step_stats *s) {
  if (c->steps_done > 0) copy3(c, c->ucont_rm1, c->ucont_o);
  copy3(c, c->ucont_o, c->ucont);
  return c->mom.scheme == CV_MOM_IMPRK ? cv_imprk_dev(c, s) : cv_rk_dev(c);
}

// the part after it: UpdatePressure and Projection (solvers.c:1066-1068)
int cv_step_post_dev(cv_ctx *c) {
  int rc = cv_update_pressure_dev(c);
  return rc ? rc : cv_projection_dev(c);
}

extern "C" int cv_flow_step(cv_ctx *c, cv_step_stats *st) {
  int rc = ready(c);
  if (rc) return rc;
  cv_step_stats s{};
  hipEvent_t *e = c->step_ev;
  CV_CHECK(hipEventRecord(e[0], c->stream));
  if ((rc = cv_step_pre_dev(c, &s))) return rc;
  CV_CHECK(hipEventRecord(e[1], c->stream));
  rc = cv_poisson_solve_dev(c, &s);
  if (rc && rc != CV_ERR_NOT_CONVERGED) return rc;
  int conv_rc = rc;
  CV_CHECK(hipEventRecord(e[2], c->stream));
  if ((rc = cv_step_post_dev(c))) return rc;
  CV_CHECK(hipEventRecord(e[3], c->stream));
  if (c->inlet || c->outlet)
    CV_CHECK(hipMemcpyAsync(c->red_host + 110, c->dsc + CV_DSC_FLUXIN, 2 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  if ((rc = cv_divergence_dev(c, &s.maxdiv, &s.maxdiv_idx))) return rc;
  if ((rc = cv_ib_dev(c))) return rc;           // solvers.c:1095-1101
  CV_CHECK(hipStreamSynchronize(c->stream));
  if (c->inlet || c->outlet) { s.flux_in = c->red_host[110]; s.flux_out = c->red_host[111]; }
  float ms[3];
  for (int q = 0; q < 3; q++) CV_CHECK(hipEventElapsedTime(&ms[q], e[q], e[q + 1]));
  s.t_mom = ms[0] * 1e-3; s.t_p = ms[1] * 1e-3; s.t_pr = ms[2] * 1e-3;
  c->last = s;
  c->steps_done++;
  if (st) *st = s;
  if (!std::isfinite(s.maxdiv)) return CV_ERR_NONFINITE;
  return conv_rc;
}

extern "C" int cv_imprk(cv_ctx *c, cv_step_stats *st) {
  int rc = ready(c);
  if (rc) return rc;
  cv_step_stats s{};
  rc = cv_imprk_dev(c, &s);
  if (rc) return rc;
  CV_CHECK(hipStreamSynchronize(c->stream));
  if (st) *st = s;
  return CV_OK;
}

extern "C" int cv_calcrhs(cv_ctx *c, double *rhs_out) {
  int rc = ready(c);
  if (rc) return rc;
  if ((rc = cv_calcrhs_dev(c))) return rc;
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rhs_out ? cv_get_field(c, CV_F_RHS, -1, rhs_out) : CV_OK;
}

// ====================================================================
// batched momentum: the level rotation, the explicit RK step and the
// ghost updates of many cases, one launch per kernel covering every case.
// A block table maps each block of a launch to (case, block of that case's
// own launch shape) and the block runs the single-case body, so each case
// gets the bits cv_rk_dev gives it alone.
// ====================================================================
namespace {

struct MCase {
  G g;
  double *uc[3], *uo[3], *urm1[3], *uca[3], *rhs[3], *rct[3];
  double *p, *nv;
  FFArgs ff;
  FM mi, mj, mk;
  C3 csi, eta, zet;
  const double *aj;
  double nu, dt, st;
  int wx, wy, wz;           // wall masks of the i, j and k faces (wall-normal flux reset)
  long n;
  int mgx, mgy, mgz;        // k-marching blocks
  int rgx, rgy, rgz, rkc;   // face-flux tiles (8 rows) and their k planes
  int nvb, npb;             // blocks of an elementwise and of a plane pass
};

#define MB_ARGS const MCase *__restrict__ cs, const uint32_t *__restrict__ tab, unsigned long long pm
#define MB_HEAD                                            \
  const uint32_t e_ = tab[blockIdx.x];                     \
  const int c = (int)(e_ >> 24), lb = (int)(e_ & 0xffffffu); \
  if (!((pm >> c) & 1ull)) return;                         \
  const MCase &C = cs[c];
#define MB_GSV GS{(long)lb * BS + threadIdx.x, (long)C.nvb * BS}
#define MB_GSP GS{(long)lb * BS + threadIdx.x, (long)C.npb * BS}

__device__ inline C3 mc3(double *const a[3]) { return C3{a[0], a[1], a[2]}; }
__device__ inline V3 mv3(double *const a[3]) { return V3{a[0], a[1], a[2]}; }

// which 0: Ucont_rm1 <- Ucont_o; 1: Ucont_o <- Ucont; 2: Rhs <- 0
__global__ void kbm_vec3(MB_ARGS, int which) {
  MB_HEAD
  const GS gs = MB_GSV;
  for (int q = 0; q < 3; q++) {
    double *d = which == 0 ? C.urm1[q] : which == 1 ? C.uo[q] : C.rhs[q];
    const double *s = which == 0 ? C.uo[q] : C.uc[q];
    for (long t = gs.t0; t < C.n; t += gs.step) d[t] = which == 2 ? 0.0 : s[t];
  }
}

__global__ void __launch_bounds__(64 * MARCH_TY) kbm_contra2cart(MB_ARGS) {
  MB_HEAD
  const int bx = lb % C.mgx, by = (lb / C.mgx) % C.mgy, bz = lb / (C.mgx * C.mgy);
  contra2cart_body(C.g, mc3(C.uc), C.csi, C.eta, C.zet, C.nv, mv3(C.uca), C.n, bx, by, bz);
}

__global__ void kbm_wall_ghost(MB_ARGS, int full) {
  MB_HEAD
  wall_ghost_body(C.g, C.g.xper, C.g.zper, full, mv3(C.uca), nullptr, MB_GSP);   // batched cases: walls and periodic faces
}

__global__ void kbm_periodic_planes(MB_ARGS, int formbcs) {
  MB_HEAD
  periodic_planes_body(C.g, mv3(C.uca), formbcs ? C.p : nullptr, formbcs ? C.nv : nullptr, MB_GSP);
}

__global__ void kbm_edge_average(MB_ARGS, int formbcs) {
  MB_HEAD
  (void)lb;
  edge_average_body(C.g, mv3(C.uca), formbcs ? C.p : nullptr);
}

__global__ void kbm_periodic_full(MB_ARGS, int axis, int formbcs) {
  MB_HEAD
  (void)lb;   // one block per case: the ring is at most 2 (my + mz) or 2 (mx + my) points
  periodic_full_body(C.g, axis, axis == 0 ? formbcs : 0, mv3(C.uca), formbcs ? C.p : nullptr,
                     formbcs ? C.nv : nullptr, GS{(long)threadIdx.x, (long)blockDim.x});
}

__global__ void __launch_bounds__(64 * 8) kbm_rc_tiled(MB_ARGS) {
  MB_HEAD
  const int bx = lb % C.rgx, by = (lb / C.rgx) % C.rgy, bz = lb / (C.rgx * C.rgy);
  rc_tiled_body<8>(C.g, mc3(C.uc), mc3(C.uca), C.nv, C.mi, C.mj, C.mk, C.nu, C.csi, C.eta, C.zet, C.aj, mv3(C.rct),
                   C.rkc, bx, by, bz);
}

__global__ void kbm_rct_periodic(MB_ARGS) {
  MB_HEAD
  rct_periodic_body(C.g, mv3(C.rct), MB_GSP);
}

__global__ void __launch_bounds__(BS) kbm_ff1_faces(MB_ARGS) {
  MB_HEAD
  const int bx = lb % C.mgx, by = (lb / C.mgx) % C.mgy, bz = lb / (C.mgx * C.mgy);
  ff1_faces_body(C.ff, mv3(C.rhs), C.n, bx, by, bz);
}

__global__ void kbm_ff1_periodic(MB_ARGS) {
  MB_HEAD
  ff1_periodic_body(C.ff, mv3(C.rhs), MB_GSP);
}

__global__ void kbm_rk_update(MB_ARGS, double alfa) {
  MB_HEAD
  const double a = alfa * C.dt * C.st;
  rk_update_body(a, mc3(C.uo), mc3(C.rhs), mv3(C.uc), C.n, MB_GSV);
}

__global__ void kbm_wall_flux_zero(MB_ARGS) {
  MB_HEAD
  wall_flux_zero_body(C.g, C.wx, C.wy, C.wz, mv3(C.uc), MB_GSP);
}

// k_rk_c2c of every case (the batched cases have only wall and periodic faces)
__global__ void __launch_bounds__(64 * MARCH_TY) kbm_rk_c2c(MB_ARGS, double alfa) {
  MB_HEAD
  const double a = alfa * C.dt * C.st;
  const int bx = lb % C.mgx, by = (lb / C.mgx) % C.mgy, bz = lb / (C.mgx * C.mgy);
  rk_c2c_body(C.g, a, mc3(C.uo), mc3(C.rhs), mv3(C.uc), C.wx, C.wy, C.wz, C.csi, C.eta, C.zet, C.nv, mv3(C.uca), bx, by,
              bz);
}

}  // namespace

// one case's IB reconstruction buffers, for the batched sweeps
struct BIbCase {
  G g;
  int n, nfx, nfy, nface, ncell, freeslip;
  double dt;
  const int *node, *face, *cell;
  const double *w, *nv;
  double *out, *p;
  double *ucat[3], *ucont[3];
  C3 csi, eta, zet;
};

// the four IB kernels of a sweep for every case of a batch: blockIdx.y is
// the case, and each item runs the body the single-case kernels run
__global__ void kbi_nodes(const BIbCase *cs, int with_p) {
  const BIbCase &C = cs[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < C.n)
    ib_node_body(q, C.node, C.w, C.freeslip, C.dt, C3{C.ucat[0], C.ucat[1], C.ucat[2]}, C.p, with_p, C.out);
}
__global__ void kbi_scatter(const BIbCase *cs, int with_p) {
  const BIbCase &C = cs[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < C.n) ib_scatter_body(q, C.node, C.out, V3{C.ucat[0], C.ucat[1], C.ucat[2]}, C.p, with_p);
}
__global__ void kbi_faces(const BIbCase *cs) {
  const BIbCase &C = cs[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < C.nface)
    ib_face_body(q, C.g, C.nfx, C.nfy, C.face, C3{C.ucat[0], C.ucat[1], C.ucat[2]}, C.csi, C.eta, C.zet,
                 V3{C.ucont[0], C.ucont[1], C.ucont[2]});
}
__global__ void kbi_c2c(const BIbCase *cs) {
  const BIbCase &C = cs[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= C.ncell) return;
  int i, j, k;
  C.g.dec(C.cell[q], i, j, k);
  c2c_cell(C.g, C3{C.ucont[0], C.ucont[1], C.ucont[2]}, C.csi, C.eta, C.zet, C.nv, V3{C.ucat[0], C.ucat[1], C.ucat[2]},
           i, j, k);
}

struct CvMomBatch {
  int n = 0;
  hipStream_t s = nullptr;
  std::vector<cv_ctx *> cs;
  MCase *d = nullptr;
  uint32_t *tab = nullptr;
  BIbCase *ib = nullptr;          // the cases with IB nodes
  int nib = 0, ib_maxn = 0, ib_maxface = 0, ib_maxcell = 0;
  long t_march = 0, t_tile = 0, t_vec = 0, t_plane = 0, t_one = 0;
  int n_march = 0, n_tile = 0, n_vec = 0, n_plane = 0, n_one = 0;
  unsigned long long all = 0, xper = 0, zper = 0, per = 0;
};

void cv_mom_batch_destroy(CvMomBatch *mb) {
  if (!mb) return;
  if (mb->d) (void)hipFree(mb->d);
  if (mb->tab) (void)hipFree(mb->tab);
  if (mb->ib) (void)hipFree(mb->ib);
  delete mb;
}

int cv_mom_batch_create(cv_ctx *const *cases, int n, hipStream_t s, CvMomBatch **out) {
  *out = nullptr;
  if (n < 1 || n > 64) return CV_ERR_ARG;
  CvMomBatch *mb = new CvMomBatch();
  mb->n = n;
  mb->s = s;
  mb->cs.assign(cases, cases + n);
  std::vector<MCase> hc(n);
  std::vector<uint32_t> tabs;
  std::vector<int> cnt[5];
  for (int c = 0; c < n; c++) {
    cv_ctx *x = cases[c];
    CvLevel &F = x->lev[x->nlev - 1];
    CvMetrics &M = x->m;
    MCase &m = hc[c];
    memset(&m, 0, sizeof(m));
    m.g = gview(x);
    for (int q = 0; q < 3; q++) {
      m.uc[q] = x->ucont[q]; m.uo[q] = x->ucont_o[q]; m.urm1[q] = x->ucont_rm1[q];
      m.uca[q] = x->ucat[q]; m.rhs[q] = x->rhs[q]; m.rct[q] = x->rct[q];
    }
    m.p = x->p;
    m.nv = F.nvert;
    m.ff = ffargs(x);
    m.mi = FM{c3(M.icsi), c3(M.ieta), c3(M.izet), M.iaj};
    m.mj = FM{c3(M.jcsi), c3(M.jeta), c3(M.jzet), M.jaj};
    m.mk = FM{c3(M.kcsi), c3(M.keta), c3(M.kzet), M.kaj};
    m.csi = c3(M.csi); m.eta = c3(M.eta); m.zet = c3(M.zet);
    m.aj = M.aj;
    m.nu = 1. / x->ren;
    m.dt = x->dt;
    m.st = x->st;
    m.wx = wall_mask(x->bctype[0], x->bctype[1]);
    m.wy = wall_mask(x->bctype[2], x->bctype[3]);
    m.wz = wall_mask(x->bctype[4], x->bctype[5]);
    m.n = F.d.n;
    const dim3 mg = march_grid(F.d.mx, F.d.my, F.d.mz);
    m.mgx = (int)mg.x; m.mgy = (int)mg.y; m.mgz = (int)mg.z;
    m.rkc = rc_kchunk();
    m.rgx = (F.d.mx + 62) / 63; m.rgy = (F.d.my - 1 + 6) / 7; m.rgz = (F.d.mz + m.rkc - 1) / m.rkc;
    m.nvb = (int)std::min<long>((F.d.n + BS - 1) / BS, 1024);
    const long planes = (long)F.d.my * F.d.mz + (long)F.d.mx * F.d.mz + (long)F.d.mx * F.d.my;
    m.npb = (int)std::min<long>((planes + BS - 1) / BS, 256);
    cnt[0].push_back(m.mgx * m.mgy * m.mgz);
    cnt[1].push_back(m.rgx * m.rgy * m.rgz);
    cnt[2].push_back(m.nvb);
    cnt[3].push_back(m.npb);
    cnt[4].push_back(1);
    mb->all |= 1ull << c;
    if (F.d.xper) mb->xper |= 1ull << c;
    if (F.d.zper) mb->zper |= 1ull << c;
  }
  mb->per = mb->xper | mb->zper;
  long *off[5] = {&mb->t_march, &mb->t_tile, &mb->t_vec, &mb->t_plane, &mb->t_one};
  int *len[5] = {&mb->n_march, &mb->n_tile, &mb->n_vec, &mb->n_plane, &mb->n_one};
  for (int k = 0; k < 5; k++) {
    *off[k] = (long)tabs.size();
    for (int c = 0; c < n; c++)
      for (int b = 0; b < cnt[k][c]; b++) tabs.push_back(((uint32_t)c << 24) | (uint32_t)b);
    *len[k] = (int)((long)tabs.size() - *off[k]);
  }
  std::vector<BIbCase> hib;
  for (int c = 0; c < n; c++) {
    cv_ctx *x = cases[c];
    if (x->ib_n <= 0) continue;
    CvMetrics &M = x->m;
    BIbCase b;
    memset(&b, 0, sizeof(b));
    b.g = gview(x);
    b.n = x->ib_n;
    b.nfx = x->ib_nface[0]; b.nfy = x->ib_nface[1];
    b.nface = x->ib_nface[0] + x->ib_nface[1] + x->ib_nface[2];
    b.ncell = x->ib_ncell;
    b.freeslip = x->ib_mode;
    b.dt = x->dt;
    b.node = x->ib_node; b.face = x->ib_face; b.cell = x->ib_cell;
    b.w = x->ib_w; b.nv = x->lev[x->nlev - 1].nvert;
    b.out = x->ib_new; b.p = x->p;
    for (int q = 0; q < 3; q++) { b.ucat[q] = x->ucat[q]; b.ucont[q] = x->ucont[q]; }
    b.csi = c3(M.csi); b.eta = c3(M.eta); b.zet = c3(M.zet);
    mb->ib_maxn = std::max(mb->ib_maxn, b.n);
    mb->ib_maxface = std::max(mb->ib_maxface, b.nface);
    mb->ib_maxcell = std::max(mb->ib_maxcell, b.ncell);
    hib.push_back(b);
  }
  mb->nib = (int)hib.size();
  hipError_t e = hipMalloc((void **)&mb->d, sizeof(MCase) * n);
  if (e == hipSuccess) e = hipMalloc((void **)&mb->tab, sizeof(uint32_t) * tabs.size());
  if (e == hipSuccess && mb->nib > 0) e = hipMalloc((void **)&mb->ib, sizeof(BIbCase) * mb->nib);
  if (e == hipSuccess && mb->nib > 0)
    e = hipMemcpy(mb->ib, hib.data(), sizeof(BIbCase) * mb->nib, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(mb->d, hc.data(), sizeof(MCase) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(mb->tab, tabs.data(), sizeof(uint32_t) * tabs.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    cv_set_error("cv_mom_batch_create: %s", hipGetErrorString(e));
    cv_mom_batch_destroy(mb);
    return CV_ERR_HIP;
  }
  *out = mb;
  return CV_OK;
}

#define MB_LAUNCH(kern, blk, t, nb, pm, ...)                                                                      \
  do {                                                                                                          \
    if ((nb) > 0 && (pm))                                                                                       \
      hipLaunchKernelGGL(kern, dim3((unsigned)(nb)), blk, 0, mb->s, (const MCase *)mb->d,                       \
                         (const uint32_t *)(mb->tab + (t)), (unsigned long long)(pm), ##__VA_ARGS__);            \
  } while (0)

// Contra2Cart (formbcs) and the ghost pass of cv_ghost_update for every
// case: the batched cases have only wall and periodic faces, so FormBCS's
// later passes would rewrite the first pass's bits (cv_ghost_update)
int cv_mom_batch_ghost(CvMomBatch *mb, int formbcs, int cart) {
  if (cart)
    MB_LAUNCH(kbm_contra2cart, dim3(64, MARCH_TY), mb->t_march, mb->n_march, mb->all);
  MB_LAUNCH(kbm_wall_ghost, dim3(BS), mb->t_plane, mb->n_plane, mb->all, formbcs ? 0 : 1);
  MB_LAUNCH(kbm_periodic_planes, dim3(BS), mb->t_plane, mb->n_plane, mb->per, formbcs);
  MB_LAUNCH(kbm_edge_average, dim3(BS), mb->t_one, mb->n_one, mb->all, formbcs);
  MB_LAUNCH(kbm_periodic_full, dim3(BS), mb->t_one, mb->n_one, mb->xper, 0, formbcs);
  MB_LAUNCH(kbm_periodic_full, dim3(BS), mb->t_one, mb->n_one, mb->zper, 2, formbcs);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// ibm_interpolation_advanced of every case with IB nodes: the six sweeps of
// cv_ib_dev_on, each kernel launched once for all those cases
int cv_mom_batch_ib(CvMomBatch *mb) {
  if (mb->nib == 0) return CV_OK;
  const unsigned nc = (unsigned)mb->nib;
  const dim3 gn((unsigned)((mb->ib_maxn + BS - 1) / BS), nc);
  const dim3 gf((unsigned)((mb->ib_maxface + BS - 1) / BS), nc);
  const dim3 gc((unsigned)((mb->ib_maxcell + BS - 1) / BS), nc);
  for (int sweep = 0; sweep < 6; sweep++) {
    const int with_p = sweep == 0;
    hipLaunchKernelGGL(kbi_nodes, gn, dim3(BS), 0, mb->s, (const BIbCase *)mb->ib, with_p);
    hipLaunchKernelGGL(kbi_scatter, gn, dim3(BS), 0, mb->s, (const BIbCase *)mb->ib, with_p);
    if (mb->ib_maxface > 0) hipLaunchKernelGGL(kbi_faces, gf, dim3(BS), 0, mb->s, (const BIbCase *)mb->ib);
    if (mb->ib_maxcell > 0) hipLaunchKernelGGL(kbi_c2c, gc, dim3(BS), 0, mb->s, (const BIbCase *)mb->ib);
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// the level rotation (rot: the cases with a step history) and the explicit
// RK step of cv_rk_dev for every case
int cv_mom_batch_rk(CvMomBatch *mb, unsigned long long rot) {
  MB_LAUNCH(kbm_vec3, dim3(BS), mb->t_vec, mb->n_vec, rot, 0);
  MB_LAUNCH(kbm_vec3, dim3(BS), mb->t_vec, mb->n_vec, mb->all, 1);
  int rc = cv_mom_batch_ghost(mb, 1, 1);
  if (rc) return rc;
  if ((rc = cv_mom_batch_ib(mb))) return rc;   // implicitsolver.c:47-51
  MB_LAUNCH(kbm_vec3, dim3(BS), mb->t_vec, mb->n_vec, mb->all, 2);
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  for (int s = 0; s < 4; s++) {
    MB_LAUNCH(kbm_rc_tiled, dim3(64, 8), mb->t_tile, mb->n_tile, mb->all);
    MB_LAUNCH(kbm_rct_periodic, dim3(BS), mb->t_plane, mb->n_plane, mb->per);
    MB_LAUNCH(kbm_ff1_faces, dim3(64, MARCH_TY), mb->t_march, mb->n_march, mb->all);
    MB_LAUNCH(kbm_ff1_periodic, dim3(BS), mb->t_plane, mb->n_plane, mb->per);
    if (rk_fused()) {   // the stage update, the wall flux reset and Contra2Cart in one k-march
      MB_LAUNCH(kbm_rk_c2c, dim3(64, MARCH_TY), mb->t_march, mb->n_march, mb->all, alfa[s]);
      if ((rc = cv_mom_batch_ghost(mb, 1, 0))) return rc;
      continue;
    }
    MB_LAUNCH(kbm_rk_update, dim3(BS), mb->t_vec, mb->n_vec, mb->all, alfa[s]);
    MB_LAUNCH(kbm_wall_flux_zero, dim3(BS), mb->t_plane, mb->n_plane, mb->all);
    if ((rc = cv_mom_batch_ghost(mb, 1, 1))) return rc;
  }
  if ((rc = cv_mom_batch_ib(mb))) return rc;   // after the stages, implicitsolver.c:87-90
  CV_CHECK(hipGetLastError());
  return CV_OK;
}
