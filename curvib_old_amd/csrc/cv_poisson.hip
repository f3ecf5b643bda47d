// cv_poisson.hip -- pressure Poisson path on gfx950.
//
//  * per-cell operator word: face enables and the cross-derivative modes of
//    PoissonLHSNew (poisson.c:600-1507), built once per nvert change
//  * matrix-free 19-point apply in flux form, FP64 and FP32
//  * geometric multigrid with the reference's transfers (MyRestriction
//    poisson.c:129-226, MyInterpolation poisson.c:228-370) and nvert
//    hierarchy (MyNvertRestriction poisson.c:1662-1773)
//  * FGMRES (right preconditioned, classical Gram-Schmidt with DGKS
//    re-orthogonalisation), FP64 outer loop, FP64 or FP32 V-cycle
//  * PoissonRHS, VolumeFlux_rev, UpdatePressure, Projection, Divergence
//
// All kernels are memory-bound stencils over i-fastest planar arrays; no
// MFMA.  Reductions are two-pass fixed-order trees (deterministic).
#include "cv_internal.h"
#include "cv_reduce.h"
#include <cmath>
#include <chrono>
#include <mutex>

#define BS 256

// out[q] = sum_b part[q*nb + b] in a fixed order (one block per q)
__device__ __forceinline__ void final_sum_body(const double *part, int nb, double *out, int q) {
  __shared__ double sh[BS / 64];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)q * nb + b];
  s = cv_block_sum(s, sh);
  if (threadIdx.x == 0) out[q] = s;
}

__global__ void cv_k_final_sum(const double *part, int nb, double *out, int nq) {
  const int q = blockIdx.x;
  if (q >= nq) return;
  final_sum_body(part, nb, out, q);
}

namespace {

struct Idx {
  int mx, my, mz;
  // storage index -> (i, j, k); storage sizes stay below 2^31 (cv_ctx_create)
  __device__ void dec(long t, int &i, int &j, int &k) const {
    unsigned u = (unsigned)t;
    i = (int)(u % (unsigned)mx);
    u /= (unsigned)mx;
    j = (int)(u % (unsigned)my);
    k = (int)(u / (unsigned)my);
  }
  __device__ long at(int i, int j, int k) const { return (long)i + (long)mx * ((long)j + (long)my * k); }
};

// (i, j, k) of a grid-stride walk t0, t0 + step, ...: one decode of t0 and
// of step, then carries per advance instead of two divisions per point
struct IdxWalk {
  int i, j, k, di, dj, dk, mx, my;
  __device__ IdxWalk(const Idx &I, long t0, long step) : mx(I.mx), my(I.my) {
    I.dec(t0, i, j, k);
    I.dec(step, di, dj, dk);
  }
  __device__ void next() {
    i += di;
    int cj = 0;
    if (i >= mx) { i -= mx; cj = 1; }
    j += dj + cj;
    int ck = 0;
    if (j >= my) { j -= my; ck = 1; }
    k += dk + ck;
  }
};

// ------------------------------------------------------------------
// cross-derivative mode (poisson.c:800-837 and siblings)
// ------------------------------------------------------------------
__device__ inline int xmode(double hi, double lo, int at_hi, int at_lo, int per) {
  if ((at_hi && !per) || hi > 0.1) {
    if (lo < 0.1 && (!at_lo || per)) return XM_LOWER;
    return XM_DROP;
  } else if ((at_lo && !per) || lo > 0.1) {
    if (hi < 0.1) return XM_UPPER;
    return XM_DROP;
  }
  return XM_CENTRAL;
}

__global__ void k_flags(Idx I, int xper, int zper, const double *__restrict__ nv, uint32_t *__restrict__ fl, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    I.dec(t, i, j, k);
    int m[3] = {I.mx, I.my, I.mz}, c[3] = {i, j, k}, per[3] = {xper, 0, zper};
    uint32_t w = 0;
    bool interior = i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1 && k > 0 && k < I.mz - 1;
    if (interior && !(nv[t] > 0.1)) w |= 1u;
    for (int d = 0; d < 3; d++) {
      // face (i,j,k) of axis d: lower cell index in [0, m_d-2], others interior
      bool valid = c[d] <= m[d] - 2;
      for (int a = 0; a < 3; a++)
        if (a != d && (c[a] < 1 || c[a] > m[a] - 2)) valid = false;
      if (!valid) continue;
      int e[3] = {0, 0, 0}; e[d] = 1;
      long up = I.at(i + e[0], j + e[1], k + e[2]);
      bool wall = !per[d] && (c[d] == 0 || c[d] == m[d] - 2);
      if (nv[t] < 0.1 && nv[up] < 0.1 && !wall) w |= 1u << (1 + 5 * d);
      int q = 0;
      for (int tt = 0; tt < 3; tt++) {
        if (tt == d) continue;
        int f[3] = {0, 0, 0}; f[tt] = 1;
        double hi = nv[I.at(i + f[0], j + f[1], k + f[2])] + nv[I.at(i + e[0] + f[0], j + e[1] + f[1], k + e[2] + f[2])];
        double lo = nv[I.at(i - f[0], j - f[1], k - f[2])] + nv[I.at(i + e[0] - f[0], j + e[1] - f[1], k + e[2] - f[2])];
        int mode = xmode(hi, lo, c[tt] == m[tt] - 2, c[tt] == 1, per[tt]);
        w |= (uint32_t)mode << (2 + 5 * d + 2 * q);
        q++;
      }
    }
    fl[t] = w;
  }
}

// ------------------------------------------------------------------
// operator view and flux-form apply
// ------------------------------------------------------------------
template <class T>
struct OpV {
  const T *g[9];
  const T *aj;
  // CV_MIXED16 per cell, g scaled by 2^e in FP16 and aj by 2^-e in FP32
  // (aj * 2^-e * sum(g * 2^e ...) = aj * sum(g ...)):
  //   hA (16 bytes): g0 g1 | g2 g6 | g7 g8 | aj
  //   hB (8 bytes):  g3 g4 | g5, the operator word (it uses 16 bits)
  // so the j-face below a tile row needs hB alone
  const uint4 *hA;
  const uint2 *hB;
  // CV_MIXED16 y-line factors per cell: yh {a in FP16 | ip * 2^e in FP16}
  // and yc {cp in FP16} (6 bytes a row); ysc = 2^-e undoes the scaling of
  // the solve
  const uint32_t *yh;
  const uint16_t *yc;
  float ysc;
  const T *dinv;
  const uint32_t *fl;
  const T *ylo, *yip, *ycp;
  int mx, my, mz, xper, zper;
};

template <class T>
__device__ inline long wrapped(const OpV<T> &A, int i, int j, int k) {
  i = cv_wrap(i, A.mx, A.xper);
  k = cv_wrap(k, A.mz, A.zper);
  return (long)i + (long)A.mx * ((long)j + (long)A.my * k);
}

// flux through face (i,j,k) of axis d (lower cell (i,j,k)); the face word w.
// X(q) returns the vector entry at storage index q.
template <class T, class XA>
__device__ inline T face_flux(const OpV<T> &A, const XA &X, int d, int i, int j, int k, uint32_t w) {
  int e[3] = {0, 0, 0}; e[d] = 1;
  const long lo = wrapped(A, i, j, k);
  const long up = wrapped(A, i + e[0], j + e[1], k + e[2]);
  const long fi = (long)i + (long)A.mx * ((long)j + (long)A.my * k);
  const T xl = X(lo), xu = X(up);
  T F = A.g[4 * d][fi] * (xu - xl);
  int q = 0;
#pragma unroll
  for (int t = 0; t < 3; t++) {
    if (t == d) continue;
    int mode = (w >> (2 + 5 * d + 2 * q)) & 3u;
    q++;
    if (mode == XM_DROP) continue;
    int f[3] = {0, 0, 0}; f[t] = 1;
    const T gdt = A.g[3 * d + t][fi];
    if (mode == XM_CENTRAL) {
      T s = X(wrapped(A, i + f[0], j + f[1], k + f[2])) + X(wrapped(A, i + e[0] + f[0], j + e[1] + f[1], k + e[2] + f[2])) -
            X(wrapped(A, i - f[0], j - f[1], k - f[2])) - X(wrapped(A, i + e[0] - f[0], j + e[1] - f[1], k + e[2] - f[2]));
      F += gdt * (T)0.25 * s;
    } else if (mode == XM_LOWER) {
      T s = xl + xu - X(wrapped(A, i - f[0], j - f[1], k - f[2])) - X(wrapped(A, i + e[0] - f[0], j + e[1] - f[1], k + e[2] - f[2]));
      F += gdt * (T)0.5 * s;
    } else {
      T s = X(wrapped(A, i + f[0], j + f[1], k + f[2])) + X(wrapped(A, i + e[0] + f[0], j + e[1] + f[1], k + e[2] + f[2])) - xl - xu;
      F += gdt * (T)0.5 * s;
    }
  }
  return F;
}

template <class T, class XA>
__device__ inline T apply_row_x(const OpV<T> &A, const XA &X, int i, int j, int k, long t) {
  const uint32_t wc = A.fl[t];
  if (!(wc & 1u)) return X(t);
  T acc = 0;
  const uint32_t wi = A.fl[t - 1], wj = A.fl[t - A.mx], wk = A.fl[t - (long)A.mx * A.my];
  if (wc & (1u << 1)) acc += face_flux(A, X, 0, i, j, k, wc);
  if (wi & (1u << 1)) acc -= face_flux(A, X, 0, i - 1, j, k, wi);
  if (wc & (1u << 6)) acc += face_flux(A, X, 1, i, j, k, wc);
  if (wj & (1u << 6)) acc -= face_flux(A, X, 1, i, j - 1, k, wj);
  if (wc & (1u << 11)) acc += face_flux(A, X, 2, i, j, k, wc);
  if (wk & (1u << 11)) acc -= face_flux(A, X, 2, i, j, k - 1, wk);
  return -A.aj[t] * acc;
}

template <class T>
struct PtrX {
  const T *__restrict__ p;
  __device__ T operator()(long q) const { return p[q]; }
};
template <class T>
struct UnitX {          // the unit vector at storage index q0
  long q0;
  __device__ T operator()(long q) const { return q == q0 ? (T)1 : (T)0; }
};

template <class T>
__device__ inline T apply_row(const OpV<T> &A, const T *__restrict__ x, int i, int j, int k, long t) {
  return apply_row_x(A, PtrX<T>{x}, i, j, k, t);
}

// damped Jacobi sweep: xo = xi + w*D^-1*(b - A xi); zero_guess: xi == 0
template <class T>
__global__ void __launch_bounds__(BS) k_jacobi(OpV<T> A, const T *__restrict__ xi, const T *__restrict__ b, T *__restrict__ xo,
                                               T omega, int zero_guess, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    if (zero_guess) { xo[t] = omega * A.dinv[t] * b[t]; continue; }
    int i = (int)(t % A.mx), j = (int)((t / A.mx) % A.my), k = (int)(t / ((long)A.mx * A.my));
    T r = b[t] - apply_row(A, xi, i, j, k, t);
    xo[t] = xi[t] + omega * A.dinv[t] * r;
  }
}

// ------------------------------------------------------------------
// 2.5D-tiled flux-form stencil.  A block of TW x TY threads (TW = 64 or
// 32 lanes per tile row) owns TW-1 cells along i (lane L <-> cell i0-1+L;
// lane 0 only supplies the flux of face i0-1) and TY rows along j, and
// marches k over one chunk.  The x
// planes k-1, k, k+1 of the tile (+1 halo, periodic wrap applied at load)
// live in LDS (four rotating slots); every face flux is formed once: the
// lower i-face comes from the neighbouring lane (shuffle), the lower j-face
// from the row below (LDS, double buffered), the lower k-face from the
// previous plane (register).  One barrier per plane; the x plane two
// ahead and the per-cell coefficients one plane ahead are loaded into
// registers before the current plane is computed.  The arithmetic per face
// and per row is that of face_flux/apply_row_x.
//   MODE 0: y = A x      MODE 1: y = b - A x
//   MODE 2: y = x + omega * dinv * (b - A x)   (damped Jacobi sweep)
// ------------------------------------------------------------------

// one face flux from its two cells (xl, xu) and the four neighbours of
// each tangential axis: (+f, +e+f, -f, +e-f)
template <class T>
__device__ inline T tile_flux(T xl, T xu, T gdd, uint32_t w, int sh, T g0, T a0, T b0, T c0, T d0, T g1, T a1, T b1,
                              T c1, T d1) {
  T F = gdd * (xu - xl);
  const int m0 = (w >> sh) & 3u, m1 = (w >> (sh + 2)) & 3u;
  if (m0 == XM_CENTRAL) F += g0 * (T)0.25 * (a0 + b0 - c0 - d0);
  else if (m0 == XM_LOWER) F += g0 * (T)0.5 * (xl + xu - c0 - d0);
  else if (m0 == XM_UPPER) F += g0 * (T)0.5 * (a0 + b0 - xl - xu);
  if (m1 == XM_CENTRAL) F += g1 * (T)0.25 * (a1 + b1 - c1 - d1);
  else if (m1 == XM_LOWER) F += g1 * (T)0.5 * (xl + xu - c1 - d1);
  else if (m1 == XM_UPPER) F += g1 * (T)0.5 * (a1 + b1 - xl - xu);
  return F;
}

// k-face flux between planes P (k) and Q (k+1) at LDS column c, row r;
// g6, g7, g8 of the face; ST_TC: the LDS row stride (tile width + 2)
template <class T, int ST_TC>
__device__ inline T tile_flux_k(const T *P, const T *Q, int c, int r, uint32_t w, T g6, T g7, T g8) {
  if (!(w & (1u << 11))) return (T)0;
  const int o = r * ST_TC + c;
  return tile_flux<T>(P[o], Q[o], g8, w, 12, g6, P[o + 1], Q[o + 1], P[o - 1], Q[o - 1],
                      g7, P[o + ST_TC], Q[o + ST_TC], P[o - ST_TC], Q[o - ST_TC]);
}
// j-face flux between rows r and r+1 of plane P (M = k-1, Q = k+1)
template <class T, int ST_TC>
__device__ inline T tile_flux_j(const T *M, const T *P, const T *Q, int c, int r, uint32_t w, T g3, T g4, T g5) {
  if (!(w & (1u << 6))) return (T)0;
  const int o = r * ST_TC + c, u = o + ST_TC;
  return tile_flux<T>(P[o], P[u], g4, w, 7, g3, P[o + 1], P[u + 1], P[o - 1], P[u - 1],
                      g5, Q[o], Q[u], M[o], M[u]);
}
// i-face flux between columns c and c+1 of plane P
template <class T, int ST_TC>
__device__ inline T tile_flux_i(const T *M, const T *P, const T *Q, int c, int r, uint32_t w, T g0, T g1, T g2) {
  if (!(w & (1u << 1))) return (T)0;
  const int o = r * ST_TC + c;
  return tile_flux<T>(P[o], P[o + 1], g0, w, 2, g1, P[o + ST_TC], P[o + ST_TC + 1], P[o - ST_TC], P[o - ST_TC + 1],
                      g2, Q[o], Q[o + 1], M[o], M[o + 1]);
}

// FP16 halves of packed coefficient words, widened to FP32 (exact)
__device__ inline float half_lo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu)); }
__device__ inline float half_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)); }

// operator word and face coefficients of one cell, loaded a plane ahead;
// in FP32 (where registers allow it) also aj, b and dinv
template <class T>
struct StCell {
  uint32_t w, wb;        // operator words of the cell and (row 0) of the cell below the tile
  T g[9], gb[3];
  T aj, b, dv;
  // FP16 packs as loaded (HG): decoded into w, g, aj, wb, gb only when the
  // plane is computed, so the prefetch a plane ahead stays in flight
  uint4 ha;
  uint2 hb, hbb;
};

// TY: rows per block (block TW x TY); PFB: prefetch the row-0 face below
// the tile with the cells (keeps the row-0 wave off the barrier's critical
// path at the cost of registers); HG: face coefficients from the FP16
// packs instead of g, aj; TW: lanes per tile row, 64 (one wavefront per
// row) or 32 (two rows per wavefront; the i-face shuffle stays within a
// row); remap: XCD-aware tile order
// one tile (bx, by, bz) of the stencil: (TW-1) x TY cells of kchunk planes
// TX: element type of x (FP32 preconditioned vectors under the FP64 apply
// of the mixed solve; widened exactly on load)
template <class T, int MODE, int TY, bool PFB, bool HG = false, int TW = 64, class TX = T>
__device__ __forceinline__ void stencil_tile(const OpV<T> &A, const TX *__restrict__ x, const T *__restrict__ b,
                                    T *__restrict__ y, const T *__restrict__ dinv, T omega, int kchunk, int bx,
                                    int by, int bz, int color = -1) {
  static_assert(TW == 64 || TW == 32, "tile rows of 64 or 32 lanes");
  constexpr int ST_TC = TW + 2;
  constexpr int TR = TY + 2, NT = TR * ST_TC, NTH = TW * TY;
  static_assert(2 * NTH >= NT, "two tile entries per thread must cover a plane");
  __shared__ T X[4][NT];
  __shared__ T FJ[2][TY][TW];
  const int L = threadIdx.x, R = threadIdx.y;
  const int tid = R * TW + L;
  const int mx = A.mx, my = A.my, mz = A.mz;
  const int i0 = (TW - 1) * bx, j0 = TY * by;
  const int k0 = kchunk * bz, k1 = min(k0 + kchunk, mz);
  const int gi = i0 - 1 + L, gj = j0 + R;
  const bool in_i = gi >= 0 && gi < mx, in_j = gj < my;
  const bool cell = in_i && in_j;
  const bool out = L >= 1 && cell;
  const bool row0 = out && R == 0 && gj >= 1;
  const long pl = (long)mx * my;
  const int c = L + 1, r = R + 1;
  const long tij = (long)gi + (long)mx * gj;

  // this thread's share of a tile plane: entries tid and e1 = tid + NTH
  const int e1 = tid + NTH;
  long off0 = 0, off1 = 0;
  bool ok0, ok1;
  {
    int rr = tid / ST_TC, cc = tid - rr * ST_TC;
    int ii = i0 - 2 + cc, jj = j0 - 1 + rr;
    ok0 = tid < NT && ii >= 0 && ii < mx && jj >= 0 && jj < my;
    if (ok0) off0 = (long)mx * jj + cv_wrap(ii, mx, A.xper);
    rr = e1 / ST_TC; cc = e1 - rr * ST_TC;
    ii = i0 - 2 + cc; jj = j0 - 1 + rr;
    ok1 = e1 < NT && ii >= 0 && ii < mx && jj >= 0 && jj < my;
    if (ok1) off1 = (long)mx * jj + cv_wrap(ii, mx, A.xper);
  }
  // x values travel in their own type until they are stored to LDS, so a
  // narrower x (FP32 under the FP64 apply) is widened only when its
  // prefetch is consumed, not when it is issued
  auto fetch_x = [&](int gk, TX &v0, TX &v1) {
    const bool kin = gk >= 0 && gk < mz;
    const long kb = kin ? pl * cv_wrap(gk, mz, A.zper) : 0;
    v0 = (kin && ok0) ? x[kb + off0] : (TX)0;
    v1 = (kin && ok1) ? x[kb + off1] : (TX)0;
  };
  auto store_x = [&](int gk, TX v0, TX v1) {
    T *S = X[(gk + 4) & 3];
    if (tid < NT) S[tid] = (T)v0;
    if (e1 < NT) S[e1] = (T)v1;
  };
  constexpr bool PF = sizeof(T) == 4;      // prefetch aj, b, dinv with the cell
  auto gco = [&](int q, long t) -> T {
    if constexpr (HG) {
      if (q >= 3 && q <= 5) {
        const uint2 hb = A.hB[t];
        return q == 3 ? half_lo(hb.x) : q == 4 ? half_hi(hb.x) : half_lo(hb.y);
      }
      const uint4 ha = A.hA[t];
      return q == 0 ? half_lo(ha.x) : q == 1 ? half_hi(ha.x) : q == 2 ? half_lo(ha.y)
           : q == 6 ? half_hi(ha.y) : q == 7 ? half_lo(ha.z) : half_hi(ha.z);
    } else {
      return A.g[q][t];
    }
  };
  auto ajo = [&](long t) -> T {
    if constexpr (HG) return __uint_as_float(A.hA[t].w);
    else return A.aj[t];
  };
  auto wo = [&](long t) -> uint32_t {
    if constexpr (HG) return A.hB[t].y >> 16;
    else return A.fl[t];
  };
  auto fetch_cell = [&](int gk, StCell<T> &C) {
    C.w = 0u;
#pragma unroll
    for (int q = 0; q < 9; q++) C.g[q] = (T)0;
    C.aj = C.b = C.dv = (T)0;
    C.wb = 0u;
    C.gb[0] = C.gb[1] = C.gb[2] = (T)0;
    if constexpr (HG) {
      C.ha = make_uint4(0u, 0u, 0u, 0u);
      C.hb = C.hbb = make_uint2(0u, 0u);
    }
    if (!cell) return;
    const long t = tij + pl * gk;
    if constexpr (HG) {
      C.ha = A.hA[t];
      C.hb = A.hB[t];
    } else {
      C.w = A.fl[t];
#pragma unroll
      for (int q = 0; q < 9; q++) C.g[q] = A.g[q][t];
    }
    if (PF && out) {
      if constexpr (!HG) C.aj = A.aj[t];
      if (MODE >= 1) C.b = b[t];
      if (MODE == 2) C.dv = dinv[t];
    }
    if (PFB && row0) {
      if constexpr (HG) {
        C.hbb = A.hB[t - mx];
      } else {
        C.wb = A.fl[t - mx];
        C.gb[0] = A.g[3][t - mx]; C.gb[1] = A.g[4][t - mx]; C.gb[2] = A.g[5][t - mx];
      }
    }
  };

  // the FP16 packs of a fetched cell into its operands (zero where the cell
  // lies outside the grid: the packs were cleared then)
  auto decode_cell = [&](StCell<T> &C) {
    if constexpr (HG) {
      const uint4 ha = C.ha;
      const uint2 hb = C.hb, hbb = C.hbb;
      C.w = hb.y >> 16;
      C.g[0] = half_lo(ha.x); C.g[1] = half_hi(ha.x); C.g[2] = half_lo(ha.y);
      C.g[3] = half_lo(hb.x); C.g[4] = half_hi(hb.x); C.g[5] = half_lo(hb.y);
      C.g[6] = half_hi(ha.y); C.g[7] = half_lo(ha.z); C.g[8] = half_hi(ha.z);
      C.aj = out ? __uint_as_float(ha.w) : (T)0;
      if (PFB && row0) {
        C.wb = hbb.y >> 16;
        C.gb[0] = half_lo(hbb.x); C.gb[1] = half_hi(hbb.x); C.gb[2] = half_lo(hbb.y);
      }
    }
  };

  // prologue: planes k0-1 .. k0+1 in LDS, plane k0+2 and the cells of k0 in
  // registers.  The chunk reads planes k0-1 .. k1 only (plane gk+1 is the
  // last neighbour of plane gk), so planes past k1 are neither loaded nor
  // stored: the next chunk loads them itself
  TX xa, xb;
  fetch_x(k0 - 1, xa, xb); store_x(k0 - 1, xa, xb);
  fetch_x(k0, xa, xb); store_x(k0, xa, xb);
  fetch_x(k0 + 1, xa, xb); store_x(k0 + 1, xa, xb);
  TX xn0 = (TX)0, xn1 = (TX)0;
  if (k0 + 2 <= k1) fetch_x(k0 + 2, xn0, xn1);
  StCell<T> cur;
  fetch_cell(k0, cur);
  __syncthreads();
  T fk_lo = (T)0;
  if (out && k0 >= 1) {
    const long t = tij + pl * (k0 - 1);
    fk_lo = tile_flux_k<T, ST_TC>(X[(k0 - 1 + 4) & 3], X[(k0 + 4) & 3], c, r, wo(t), gco(6, t), gco(7, t), gco(8, t));
  }
  for (int gk = k0; gk < k1; gk++) {
    TX xq0 = (TX)0, xq1 = (TX)0;
    if (gk + 3 <= k1) fetch_x(gk + 3, xq0, xq1);
    StCell<T> nxt;
    if (gk + 1 < k1) fetch_cell(gk + 1, nxt);
    const long t = tij + pl * gk;
    const T *Pm = X[(gk - 1 + 4) & 3], *P = X[(gk + 4) & 3], *Q = X[(gk + 1 + 4) & 3];
    decode_cell(cur);
    const uint32_t w = cur.w;
    const T fi = cell ? tile_flux_i<T, ST_TC>(Pm, P, Q, c, r, w, cur.g[0], cur.g[1], cur.g[2]) : (T)0;
    T fj = (T)0, fk = (T)0, fj_lo = (T)0;
    if (out) {
      fj = tile_flux_j<T, ST_TC>(Pm, P, Q, c, r, w, cur.g[3], cur.g[4], cur.g[5]);
      fk = tile_flux_k<T, ST_TC>(P, Q, c, r, w, cur.g[6], cur.g[7], cur.g[8]);
      if (row0) {
        if (PFB) {
          fj_lo = tile_flux_j<T, ST_TC>(Pm, P, Q, c, r - 1, cur.wb, cur.gb[0], cur.gb[1], cur.gb[2]);
        } else {
          const long tb = t - mx;
          fj_lo = tile_flux_j<T, ST_TC>(Pm, P, Q, c, r - 1, wo(tb), gco(3, tb), gco(4, tb), gco(5, tb));
        }
      }
    }
    // operands of the output (FP64: loaded here, before the barrier)
    T ajc = cur.aj, bc = cur.b, dvc = cur.dv;
    if (!PF && out) {
      ajc = ajo(t);
      if (MODE >= 1) bc = b[t];
      if (MODE == 2) dvc = dinv[t];
    }
    const T xc = P[r * ST_TC + c];
    FJ[gk & 1][R][L] = fj;
    if (gk + 2 <= k1) store_x(gk + 2, xn0, xn1);
    const T fi_lo = __shfl_up(fi, 1, TW);
    __syncthreads();
    if (R > 0) fj_lo = FJ[gk & 1][R - 1][L];
    if (out) {
      T res;
      // MODE 2 with a colour: cells of the other (i+j+k) parity keep x
      const bool keep = MODE == 2 && color >= 0 && ((gi + gj + gk) & 1) != color;
      if (!(w & 1u)) {
        const T xv = (T)x[t];
        if (MODE == 0) res = xv;
        else if (MODE == 1) res = bc - xv;
        else res = keep ? xv : xv + omega * dvc * (bc - xv);
      } else if (keep) {
        res = xc;
      } else {
        T acc = 0;
        acc += fi; acc -= fi_lo;
        acc += fj; acc -= fj_lo;
        acc += fk; acc -= fk_lo;
        const T ax = -ajc * acc;
        if (MODE == 0) res = ax;
        else if (MODE == 1) res = bc - ax;
        else res = xc + omega * dvc * (bc - ax);
      }
      y[t] = res;
    }
    fk_lo = fk;
    cur = nxt;
    xn0 = xq0; xn1 = xq1;
  }
}

template <class T, int MODE, int TY, bool PFB, bool HG = false, int TW = 64, class TX = T>
__global__ void __launch_bounds__(TW * TY) k_stencil(OpV<T> A, const TX *__restrict__ x, const T *__restrict__ b,
                                                   T *__restrict__ y, const T *__restrict__ dinv, T omega, int kchunk,
                                                   int remap, int color) {
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if (remap == 1) cv_xcd_tile(bx, by, bz);
  else if (remap > 1) cv_xcd_runs(remap, bx, by, bz);
  stencil_tile<T, MODE, TY, PFB, HG, TW, TX>(A, x, b, y, dinv, omega, kchunk, bx, by, bz, color);
}

// ------------------------------------------------------------------
// y-line (wall-normal) smoother: per row the coefficients of x[c-e_j],
// x[c], x[c+e_j] taken from the operator itself, the Thomas factors of
// every (i,k) line, and a fused residual + line-solve sweep
// ------------------------------------------------------------------
__global__ void k_linecoef(OpV<double> A, double *lo, double *dg, double *up, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i = (int)(t % A.mx), j = (int)((t / A.mx) % A.my), k = (int)(t / ((long)A.mx * A.my));
    if (!(A.fl[t] & 1u)) { lo[t] = 0.0; dg[t] = 1.0; up[t] = 0.0; continue; }
    lo[t] = apply_row_x(A, UnitX<double>{t - A.mx}, i, j, k, t);
    dg[t] = apply_row_x(A, UnitX<double>{t}, i, j, k, t);
    up[t] = apply_row_x(A, UnitX<double>{t + A.mx}, i, j, k, t);
  }
}

// Thomas factors along j=1..my-2 of every (i,k) line: ip = 1/pivot,
// cp = up/pivot, and lo replaced by a = -lo/pivot; rows j=0 and my-1 are
// identity rows
__global__ void k_linefactor(int mx, int my, int mz, double *lo, const double *dg, const double *up,
                             double *ip, double *cp, double *dinv) {
  const long nl = (long)mx * mz;
  for (long ln = blockIdx.x * (long)blockDim.x + threadIdx.x; ln < nl; ln += (long)gridDim.x * blockDim.x) {
    const int i = (int)(ln % mx), k = (int)(ln / mx);
    const long base = (long)i + (long)mx * my * k;
    double cprev = 0.0;
    for (int j = 1; j < my - 1; j++) {
      const long t = base + (long)mx * j;
      double piv = dg[t] - lo[t] * cprev;
      double r = piv != 0.0 ? 1.0 / piv : 1.0;
      ip[t] = r;
      lo[t] = -lo[t] * r;      // stored as the forward multiplier a = -lo/pivot
      cprev = up[t] * r;
      cp[t] = cprev;
      dinv[t] = dg[t] != 0.0 ? 1.0 / dg[t] : 1.0;
    }
    ip[base] = 1.0; cp[base] = 0.0; dinv[base] = 1.0;
    ip[base + (long)mx * (my - 1)] = 1.0; cp[base + (long)mx * (my - 1)] = 0.0; dinv[base + (long)mx * (my - 1)] = 1.0;
  }
}

// chunked line solve for lines longer than the register-resident kernel
// takes (lo holds a = -lo/pivot): block (64 lines in i) x (YL_NC chunks along j); the
// two recurrences of the factored Thomas solve (forward y_j = b_j + a_j
// y_{j-1}, backward e_j = y_j - cp_j e_{j+1}) are composed per chunk as
// affine maps, the chunk carries resolved through LDS, then each chunk is
// swept again from its true carry.  r: residual (or b for a zero guess).
#define YL_NC 4
template <class T>
__global__ void __launch_bounds__(64 * YL_NC) k_yline_solve(int mx, int my, int mz, const T *r,
                                                           const T *__restrict__ lo, const T *__restrict__ ip,
                                                           const T *__restrict__ cp, const T *__restrict__ xi,
                                                           T *__restrict__ xo, T *tmp, T omega,
                                                           int zero_guess) {
  __shared__ T SA[YL_NC][64], SB[YL_NC][64];
  const int ti = threadIdx.x, tc = threadIdx.y;
  const int i = blockIdx.x * 64 + ti, k = blockIdx.y;
  const bool act = i < mx;
  const long sj = mx;
  const long base = (long)i + (long)mx * my * k;
  const int nr = my - 2, cs = (nr + YL_NC - 1) / YL_NC;
  const int j0 = 1 + tc * cs, j1 = min(1 + (tc + 1) * cs, my - 1);
  // forward: composite of the chunk
  T Aa = 1, Bb = 0;
  if (act)
    for (int j = j0; j < j1; j++) {
      const long t = base + sj * j;
      const T a = lo[t], bb = r[t] * ip[t];
      Aa = a * Aa;
      Bb = a * Bb + bb;
    }
  SA[tc][ti] = Aa; SB[tc][ti] = Bb;
  __syncthreads();
  T y = 0;
  for (int q = 0; q < tc; q++) y = SA[q][ti] * y + SB[q][ti];
  if (act)
    for (int j = j0; j < j1; j++) {
      const long t = base + sj * j;
      y = r[t] * ip[t] + lo[t] * y;
      tmp[t] = y;
    }
  __syncthreads();
  // backward: composite of the chunk, processed from its top row down
  Aa = 1; Bb = 0;
  if (act)
    for (int j = j1 - 1; j >= j0; j--) {
      const long t = base + sj * j;
      const T a = -cp[t];
      Aa = a * Aa;
      Bb = a * Bb + tmp[t];
    }
  SA[tc][ti] = Aa; SB[tc][ti] = Bb;
  __syncthreads();
  T e = 0;
  for (int q = YL_NC - 1; q > tc; q--) e = SA[q][ti] * e + SB[q][ti];
  if (act) {
    for (int j = j1 - 1; j >= j0; j--) {
      const long t = base + sj * j;
      e = tmp[t] - cp[t] * e;
      xo[t] = (zero_guess ? (T)0 : xi[t]) + omega * e;
    }
    if (tc == 0) xo[base] = zero_guess ? (T)0 : xi[base];
    if (tc == YL_NC - 1) xo[base + sj * (my - 1)] = zero_guess ? (T)0 : xi[base + sj * (my - 1)];
  }
}

// register-resident line solve: block 64 lanes (i) x NC chunks along j,
// each thread keeps the CS rows of its chunk in registers, so r, the line
// factors and x are each read once and x written once.  Forward recurrence
// y_j = r_j ip_j + a_j y_{j-1} (a = -lo ip, stored), backward
// e_j = y_j - cp_j e_{j+1}; chunk carries resolved through LDS.
// lines lg*64 .. lg*64+63 of the register-resident line solve, counted
// i fastest over the (i, k) pairs of the level (a group may span two k
// planes, so no lane idles where mx is not a multiple of 64); the chunks
// beyond the line (blockDim.y larger than the line needs) are identity
// maps, so the result does not depend on blockDim.y
// HL: the factors from the FP16 packs yh and yc (a 4-byte and a 2-byte
// load per row, cp kept in registers for the backward pass); the solution
// then carries 2^e and the caller passes omega * 2^-e
// PX: x is loaded with r at the start (each thread reads and writes only
// its own rows, so xi may be xo), which takes one dependent memory round
// trip off a launch of few blocks at the cost of CS registers (launches of
// more blocks than CUs keep PX off)
template <class T, int CS, bool HL = false, bool PX = false>
__device__ __forceinline__ void yline_reg_tile(int mx, int my, int mz, const T *r, const T *__restrict__ ya,
                                      const T *__restrict__ ip, const T *__restrict__ cp, const T *xi, T *xo,
                                      T omega, int zero_guess, int lg, const uint32_t *__restrict__ yh = nullptr,
                                      const uint16_t *__restrict__ yc = nullptr) {
  __shared__ T SA[16][64], SB[16][64];
  const int ti = threadIdx.x, tc = threadIdx.y, nc = blockDim.y;
  const long q0 = (long)lg * 64 + ti;
  const bool act = q0 < (long)mx * mz;
  const int i = act ? (int)(q0 % mx) : 0, k = act ? (int)(q0 / mx) : 0;
  const long sj = mx;
  const long base = (long)i + (long)mx * my * k;
  const int j0 = 1 + tc * CS;
  const int nrow = act ? max(0, min(CS, my - 1 - j0)) : 0;
  T av[CS], yv[CS], cv[HL ? CS : 1], xv[PX ? CS : 1];
  if constexpr (PX) {
#pragma unroll
    for (int q = 0; q < CS; q++)
      xv[q] = q < nrow && !zero_guess ? xi[base + sj * (j0 + q)] : (T)0;
  }
  // forward composite of the chunk
  T Aa = 1, Bb = 0;
#pragma unroll
  for (int q = 0; q < CS; q++) {
    if (q < nrow) {
      const long t = base + sj * (j0 + q);
      if constexpr (HL) {
        const uint32_t w = yh[t];
        av[q] = half_lo(w);
        yv[q] = r[t] * half_hi(w);
        cv[q] = half_lo((uint32_t)yc[t]);
      } else {
        av[q] = ya[t];
        yv[q] = r[t] * ip[t];
      }
      Aa = av[q] * Aa;
      Bb = av[q] * Bb + yv[q];
    }
  }
  SA[tc][ti] = Aa; SB[tc][ti] = Bb;
  __syncthreads();
  T y = 0;
  for (int q = 0; q < tc; q++) y = SA[q][ti] * y + SB[q][ti];
#pragma unroll
  for (int q = 0; q < CS; q++)
    if (q < nrow) { y = yv[q] + av[q] * y; yv[q] = y; }
  __syncthreads();
  // backward composite, from the chunk's top row down
  Aa = 1; Bb = 0;
#pragma unroll
  for (int q = CS - 1; q >= 0; q--) {
    if (q < nrow) {
      if constexpr (HL) av[q] = -cv[q];
      else av[q] = -cp[base + sj * (j0 + q)];
      Aa = av[q] * Aa;
      Bb = av[q] * Bb + yv[q];
    }
  }
  SA[tc][ti] = Aa; SB[tc][ti] = Bb;
  __syncthreads();
  T e = 0;
  for (int q = nc - 1; q > tc; q--) e = SA[q][ti] * e + SB[q][ti];
  if (!act) return;
#pragma unroll
  for (int q = CS - 1; q >= 0; q--) {
    if (q < nrow) {
      const long t = base + sj * (j0 + q);
      e = yv[q] + av[q] * e;
      if constexpr (PX) xo[t] = xv[q] + omega * e;
      else xo[t] = (zero_guess ? (T)0 : xi[t]) + omega * e;
    }
  }
  if (tc == 0) xo[base] = zero_guess ? (T)0 : xi[base];
  if (tc == nc - 1) xo[base + sj * (my - 1)] = zero_guess ? (T)0 : xi[base + sj * (my - 1)];
}

template <class T, int CS, bool HL = false, bool PX = false>
__global__ void __launch_bounds__(1024) k_yline_reg(int mx, int my, int mz, const T *r, const T *__restrict__ ya,
                                                   const T *__restrict__ ip, const T *__restrict__ cp,
                                                   const T *xi, T *xo, T omega, int zero_guess,
                                                   const uint32_t *__restrict__ yh, const uint16_t *__restrict__ yc) {
  int lg, by, bz;
  cv_xcd_tile(lg, by, bz);
  yline_reg_tile<T, CS, HL, PX>(mx, my, mz, r, ya, ip, cp, xi, xo, omega, zero_guess, lg, yh, yc);
}

// ------------------------------------------------------------------
// nvert hierarchy (MyNvertRestriction, poisson.c:1662-1773)
// ------------------------------------------------------------------
__global__ void k_nvrestrict1(Idx C, Idx F, int isc, int jsc, int ksc, const double *nf, double *nc, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    C.dec(t, i, j, k);
    double v = 0.0;
    if (i > 0 && i < C.mx - 1 && j > 0 && j < C.my - 1 && k > 0 && k < C.mz - 1) {
      int ia = isc ? 0 : 1, ja = jsc ? 0 : 1, ka = ksc ? 0 : 1;
      int ih = isc ? i : 2 * i, jh = jsc ? j : 2 * j, kh = ksc ? k : 2 * k;
      double prod = nf[F.at(ih, jh, kh)] * nf[F.at(ih - ia, jh, kh)] * nf[F.at(ih, jh - ja, kh)] *
                    nf[F.at(ih, jh, kh - ka)] * nf[F.at(ih - ia, jh - ja, kh)] * nf[F.at(ih, jh - ja, kh - ka)] *
                    nf[F.at(ih - ia, jh, kh - ka)] * nf[F.at(ih - ia, jh - ja, kh - ka)];
      if (prod > 0.1) v = fmax(1., v);
    }
    nc[t] = v;
  }
}
__global__ void k_nvrestrict2(Idx C, const double *nin, double *nout, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    C.dec(t, i, j, k);
    double v = nin[t];
    if (i > 0 && i < C.mx - 1 && j > 0 && j < C.my - 1 && k > 0 && k < C.mz - 1 && v < 0.1) {
      if (nin[t + 1] + nin[t - 1] > 1.1 && nin[t + C.mx] + nin[t - C.mx] > 1.1 &&
          nin[t + (long)C.mx * C.my] + nin[t - (long)C.mx * C.my] > 1.1)
        v = 1.;
    }
    nout[t] = v;
  }
}

// ------------------------------------------------------------------
// transfers
// ------------------------------------------------------------------
// restriction of the coarse points t0, t0+step, ...  Fine points enter
// weighted by max(0, 1 - nvert); in FP32 that weight comes precomputed
// (wff, the same value rounded once).  The coarse point is an interior
// fluid cell iff bit 0 of its operator word is set.
template <class T>
__device__ __forceinline__ void restrict_body(Idx C, Idx F, int isc, int jsc, int ksc, const double *__restrict__ nvf,
                                     const float *__restrict__ wff, const uint32_t *__restrict__ flc,
                                     const T *__restrict__ xf, T *__restrict__ xc, long n, long t0, long step) {
  IdxWalk w(C, t0, step);
  for (long t = t0; t < n; t += step, w.next()) {
    const int i = w.i, j = w.j, k = w.k;
    T f = 0;
    if (flc[t] & 1u) {
      // 32-bit indices (storage sizes stay below 2^31): the fine point and
      // its lower neighbours along the coarsened axes at fixed strides
      const int ia = isc ? 0 : 1, ja = jsc ? 0 : F.mx, ka = ksc ? 0 : F.mx * F.my;
      const int ih = isc ? i : 2 * i, jh = jsc ? j : 2 * j, kh = ksc ? k : 2 * k;
      const int q0 = ih + F.mx * (jh + F.my * kh);
      auto RT = [&](int q) -> T {
        if constexpr (sizeof(T) == 4) return xf[q] * wff[q];
        else return xf[q] * (T)fmax(0., 1 - nvf[q]);
      };
      if (ja == 0) {   // y kept (semi-coarsening): four distinct terms, summed in the same order
        const T a = RT(q0), b = RT(q0 - ia), c = RT(q0 - ka), d = RT(q0 - ia - ka);
        f = (T)0.125 * (a + b + a + c + b + c + d + d);
      } else {
        f = (T)0.125 * (RT(q0) + RT(q0 - ia) + RT(q0 - ja) + RT(q0 - ka) + RT(q0 - ia - ja) + RT(q0 - ja - ka) +
                        RT(q0 - ia - ka) + RT(q0 - ia - ja - ka));
      }
    }
    xc[t] = f;
  }
}
template <class T>
__global__ void k_restrict(Idx C, Idx F, int isc, int jsc, int ksc, const double *__restrict__ nvf,
                           const float *__restrict__ wff, const uint32_t *__restrict__ flc, const T *__restrict__ xf,
                           T *__restrict__ xc, long n) {
  restrict_body<T>(C, F, isc, jsc, ksc, nvf, wff, flc, xf, xc, n, blockIdx.x * (long)blockDim.x + threadIdx.x,
                   (long)gridDim.x * blockDim.x);
}
__global__ void k_nvweight(const double *__restrict__ nv, float *__restrict__ w, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    w[t] = (float)fmax(0., 1 - nv[t]);
}

// xf (+)= P xc ; add=0 overwrites; fine points other than interior fluid
// cells (bit 0 of the operator word clear) take 0
template <class T>
__device__ __forceinline__ void prolong_body(Idx F, Idx C, int isc, int jsc, int ksc, const double *__restrict__ nvc,
                                    const uint32_t *__restrict__ flf, const T *__restrict__ xc, T *__restrict__ xf,
                                    int add, long n, long t0, long step) {
  IdxWalk w(F, t0, step);
  for (long t = t0; t < n; t += step, w.next()) {
    int i = w.i, j = w.j, k = w.k;
    T f = 0;
    if (i > 0 && i < F.mx - 1 && j > 0 && j < F.my - 1 && k > 0 && k < F.mz - 1) {
      int ic, jc, kc, ia, ja, ka;
      if (isc) { ic = i; ia = 0; } else { ic = (i + 1) / 2; ia = (i - 2 * ic) == 0 ? 1 : -1; if (i == 1 || i == F.mx - 2) ia = 0; }
      if (jsc) { jc = j; ja = 0; } else { jc = (j + 1) / 2; ja = (j - 2 * jc) == 0 ? 1 : -1; if (j == 1 || j == F.my - 2) ja = 0; }
      if (ksc) { kc = k; ka = 0; } else { kc = (k + 1) / 2; ka = (k - 2 * kc) == 0 ? 1 : -1; if (k == 1 || k == F.mz - 2) ka = 0; }
      // 32-bit indices (storage sizes stay below 2^31): the coarse point
      // and its neighbours at fixed strides
      const int c0 = ic + C.mx * (jc + C.my * kc);
      const int di = ia, dj = ja * C.mx, dk = ka * (C.mx * C.my);
      if (ka != 0 && nvc[c0 + dk] > 0.1) ka = 0;
      if (ja != 0 && nvc[c0 + dj] > 0.1) ja = 0;
      if (ia != 0 && nvc[c0 + di] > 0.1) ia = 0;
      const int oi = ia != 0 ? di : 0, oj = ja != 0 ? dj : 0, ok = ka != 0 ? dk : 0;
      if (oj == 0) {   // no j neighbour (y kept, or dropped): four distinct values, same expression
        const T v0 = xc[c0], vi = xc[c0 + oi], vk = xc[c0 + ok], vik = xc[c0 + oi + ok];
        f = (v0 * (T)9 + v0 * (T)3 + vi * (T)3 + vi) * (T)(3. / 64.) +
            (vk * (T)9 + vk * (T)3 + vik * (T)3 + vik) / (T)64.;
      } else {
        f = (xc[c0] * (T)9 + xc[c0 + oj] * (T)3 + xc[c0 + oi] * (T)3 + xc[c0 + oi + oj]) * (T)(3. / 64.) +
            (xc[c0 + ok] * (T)9 + xc[c0 + oj + ok] * (T)3 + xc[c0 + oi + ok] * (T)3 + xc[c0 + oi + oj + ok]) / (T)64.;
      }
    }
    if (!(flf[t] & 1u)) f = 0;
    if (add) xf[t] += f; else xf[t] = f;
  }
}
template <class T>
__global__ void k_prolong(Idx F, Idx C, int isc, int jsc, int ksc, const double *__restrict__ nvc,
                          const uint32_t *__restrict__ flf, const T *__restrict__ xc, T *__restrict__ xf, int add, long n) {
  prolong_body<T>(F, C, isc, jsc, ksc, nvc, flf, xc, xf, add, n, blockIdx.x * (long)blockDim.x + threadIdx.x,
                  (long)gridDim.x * blockDim.x);
}

// ------------------------------------------------------------------
// nullspace (PoissonNullSpaceFunction, poisson.c:411-558)
// ------------------------------------------------------------------
// partition of a reduction: block bid of nb, BS threads each
#define RED_LOOP for (long t = (long)bid * BS + threadIdx.x; t < n; t += (long)nb * BS)
// the fluid-cell sum of thread lt (of BS) of reduction block bid of nb
template <class T>
__device__ __forceinline__ double fluid_sum_thread(const T *__restrict__ x, const uint32_t *__restrict__ fl, long n,
                                                   int bid, int nb, int lt) {
  double s = 0.0;
  // four loads in flight per thread; the additions keep the order of a
  // plain strided loop, so the sum is the same bits
  const long step = (long)nb * BS;
  long t = (long)bid * BS + lt;
  for (; t + 3 * step < n; t += 4 * step) {
    const T x0 = x[t], x1 = x[t + step], x2 = x[t + 2 * step], x3 = x[t + 3 * step];
    const uint32_t f0 = fl[t], f1 = fl[t + step], f2 = fl[t + 2 * step], f3 = fl[t + 3 * step];
    if (f0 & 1u) s += (double)x0;
    if (f1 & 1u) s += (double)x1;
    if (f2 & 1u) s += (double)x2;
    if (f3 & 1u) s += (double)x3;
  }
  for (; t < n; t += step)
    if (fl[t] & 1u) s += (double)x[t];
  return s;
}
template <class T>
__device__ __forceinline__ void fluid_sum_body(const T *__restrict__ x, const uint32_t *__restrict__ fl, long n, double *part,
                                      int bid, int nb) {
  __shared__ double sh[BS / 64];
  double s = fluid_sum_thread<T>(x, fl, n, bid, nb, threadIdx.x);
  s = cv_block_sum(s, sh);
  if (threadIdx.x == 0) part[bid] = s;
}
template <class T>
__global__ void k_fluid_sum(const T *__restrict__ x, const uint32_t *__restrict__ fl, long n, double *part) {
  fluid_sum_body<T>(x, fl, n, part, blockIdx.x, gridDim.x);
}
template <class T>
__device__ __forceinline__ void fluid_shift_body(T *__restrict__ x, const uint32_t *__restrict__ fl, const double *sum,
                                        double inv_n, long n, long t0, long step) {
  const double mean = sum[0] * inv_n;
  for (long t = t0; t < n; t += step) x[t] = (fl[t] & 1u) ? (T)((double)x[t] - mean) : (T)0;
}
template <class T>
__global__ void k_fluid_shift(T *__restrict__ x, const uint32_t *__restrict__ fl, const double *sum, double inv_n, long n) {
  fluid_shift_body<T>(x, fl, sum, inv_n, n, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x);
}

// ------------------------------------------------------------------
// Krylov kernels
// ------------------------------------------------------------------
#define MAXV 33
// part[q*gridDim + b] = sum_t w[t]*V[q*stride+t] for q < NV; q == NV -> ||w||^2
template <int NV>
__device__ __forceinline__ void mdot_body(const double *__restrict__ w, const double *__restrict__ V, long stride, long n,
                                 double *part, int bid, int nb) {
  __shared__ double sh[BS / 64];
  double acc[NV + 1];
#pragma unroll
  for (int q = 0; q <= NV; q++) acc[q] = 0.0;
  RED_LOOP {
    const double wt = w[t];
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] += wt * V[q * stride + t];
    acc[NV] += wt * wt;
  }
#pragma unroll
  for (int q = 0; q <= NV; q++) {
    const double s = cv_block_sum(acc[q], sh);
    if (threadIdx.x == 0) part[(long)q * nb + bid] = s;
  }
}
template <int NV>
__global__ void __launch_bounds__(BS) k_mdot(const double *__restrict__ w, const double *__restrict__ V, long stride,
                                             long n, double *part) {
  mdot_body<NV>(w, V, stride, n, part, blockIdx.x, gridDim.x);
}

// w -= sum_q h[q] V_q ; part[b] = partial ||w||^2
template <int NV>
__device__ __forceinline__ void orth_body(double *__restrict__ w, const double *__restrict__ V, long stride,
                                 const double *__restrict__ h, long n, double *part, int bid, int nb) {
  __shared__ double sh[BS / 64];
  double hq[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) hq[q] = h[q];
  double s = 0.0;
  RED_LOOP {
    double v = w[t];
#pragma unroll
    for (int q = 0; q < NV; q++) v -= hq[q] * V[q * stride + t];
    w[t] = v;
    s += v * v;
  }
  s = cv_block_sum(s, sh);
  if (threadIdx.x == 0) part[bid] = s;
}
template <int NV>
__global__ void __launch_bounds__(BS) k_orth_update(double *__restrict__ w, const double *__restrict__ V, long stride,
                                                    const double *__restrict__ h, long n, double *part) {
  orth_body<NV>(w, V, stride, h, n, part, blockIdx.x, gridDim.x);
}

typedef void (*MdotFn)(const double *, const double *, long, long, double *);
typedef void (*OrthFn)(double *, const double *, long, const double *, long, double *);
#define CV_NV_LIST(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
  X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)
#define CV_MDOT_ENTRY(N) k_mdot<N>,
#define CV_ORTH_ENTRY(N) k_orth_update<N>,
static const MdotFn mdot_tbl[] = {CV_NV_LIST(CV_MDOT_ENTRY)};
static const OrthFn orth_tbl[] = {CV_NV_LIST(CV_ORTH_ENTRY)};

__device__ __forceinline__ void scale_to_body(double *__restrict__ dst, const double *__restrict__ src, double a, long n,
                                     long t0, long step) {
  for (long t = t0; t < n; t += step) dst[t] = src[t] * a;
}
__global__ void k_scale_to(double *__restrict__ dst, const double *__restrict__ src, const double *inv, long n) {
  scale_to_body(dst, src, inv[0], n, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x);
}
// the same, also writing the FP32 copy the mixed V-cycle takes as its
// right-hand side (the conversion it would otherwise make on entry)
__global__ void k_scale_to_f32(double *__restrict__ dst, const double *__restrict__ src, const double *inv,
                               float *__restrict__ f32, long n) {
  const double a = inv[0];
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const double v = src[t] * a;
    dst[t] = v;
    f32[t] = (float)v;
  }
}

// x += sum_q y[q] Z_q (Z in FP64, or the FP32 preconditioned vectors of
// the mixed solve, widened exactly)
template <class TZ>
__device__ __forceinline__ void xupdate_body(double *__restrict__ x, const TZ *__restrict__ Z, long stride,
                                    const double *__restrict__ y, int m, long n, long t0, long step) {
  __shared__ double ys[MAXV];
  if (threadIdx.x < m) ys[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  for (long t = t0; t < n; t += step) {
    double v = x[t];
    for (int q = 0; q < m; q++) v += ys[q] * (double)Z[q * stride + t];
    x[t] = v;
  }
}
template <class TZ>
__global__ void k_xupdate(double *__restrict__ x, const TZ *__restrict__ Z, long stride, const double *__restrict__ y,
                          int m, long n) {
  xupdate_body<TZ>(x, Z, stride, y, m, n, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x);
}
template <class T>
__global__ void k_copy_t(T *__restrict__ dst, const T *__restrict__ src, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = src[t];
}

__device__ __forceinline__ void sumsq_body(const double *__restrict__ x, long n, double *part, int bid, int nb) {
  __shared__ double sh[BS / 64];
  double s = 0.0;
  RED_LOOP s += x[t] * x[t];
  s = cv_block_sum(s, sh);
  if (threadIdx.x == 0) part[bid] = s;
}
__global__ void k_sumsq(const double *__restrict__ x, long n, double *part) { sumsq_body(x, n, part, blockIdx.x, gridDim.x); }

__global__ void k_d2f(float *dst, const double *src, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = (float)src[t];
}
template <class T>
__global__ void k_zero(T *dst, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = (T)0;
}

}  // namespace

// ====================================================================
// host side
// ====================================================================
static inline unsigned nblk(long n) {
  long b = (n + BS - 1) / BS;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

template <class T>
static OpV<T> opview(cv_ctx *c, int l);
template <>
OpV<double> opview<double>(cv_ctx *c, int l) {
  CvLevel &L = c->lev[l];
  OpV<double> A;
  for (int q = 0; q < 9; q++) A.g[q] = L.g[q];
  A.hA = nullptr; A.hB = nullptr; A.yh = nullptr; A.yc = nullptr; A.ysc = 1.0f;
  A.aj = L.aj; A.dinv = L.dinv; A.fl = L.flags;
  A.ylo = L.ylo; A.yip = L.yip; A.ycp = L.ycp;
  A.mx = L.d.mx; A.my = L.d.my; A.mz = L.d.mz; A.xper = L.d.xper; A.zper = L.d.zper;
  return A;
}
template <>
OpV<float> opview<float>(cv_ctx *c, int l) {
  CvLevel &L = c->lev[l];
  OpV<float> A;
  for (int q = 0; q < 9; q++) A.g[q] = L.gf[q];
  A.hA = L.gh ? (const uint4 *)L.gh : nullptr;
  A.hB = L.gh ? (const uint2 *)((const char *)L.gh + 16 * L.d.n) : nullptr;
  A.yh = L.gh ? (const uint32_t *)((const char *)L.gh + 24 * L.d.n) : nullptr;
  A.yc = L.gh ? (const uint16_t *)((const char *)L.gh + 28 * L.d.n) : nullptr;
  A.ysc = L.gh ? (float)ldexp(1.0, -L.yexp) : 1.0f;
  A.aj = L.ajf; A.dinv = L.dinvf; A.fl = L.flags;
  A.ylo = L.ylof; A.yip = L.yipf; A.ycp = L.ycpf;
  A.mx = L.d.mx; A.my = L.d.my; A.mz = L.d.mz; A.xper = L.d.xper; A.zper = L.d.zper;
  return A;
}
static Idx idx_of(const CvLevel &L) { return Idx{L.d.mx, L.d.my, L.d.mz}; }

// blocks of a Poisson-solve reduction over n points: about 8 points per
// thread, at most CV_RED_GRID; the same partition serves one case alone and
// in a batch, so the sums agree bit for bit
static int nred(long n) {
  long b = (n + (long)BS * 8 - 1) / ((long)BS * 8);
  if (b > CV_RED_GRID) b = CV_RED_GRID;
  if (b < 1) b = 1;
  return (int)b;
}

// k extent of one stencil block (tunable through CV_STENCIL_KCHUNK)
static int stencil_kchunk(int mz) {
  static int kc = -1;
  if (kc < 0) {
    const char *e = getenv("CV_STENCIL_KCHUNK");
    kc = e ? std::max(1, atoi(e)) : 0;
  }
  // default: 8 planes on deep levels (CV_STENCIL_KCHUNK_DEEP), 6 on
  // shallower ones (more blocks); the fastest measured at 512x256x256 (6
  // against 8 on the middle tier: 0.3 ms per step; 8 against 12 on the deep
  // tier: 0.35 ms, once a chunk stopped loading planes past its end).
  // Levels of few planes (mz < 64)
  // take CV_STENCIL_KCHUNK_SMALL planes per block: their launches are short
  // and latency-bound, so a shorter plane march can pay
  static int ks = -1, km = -1, kd = -1;
  if (ks < 0) {
    const char *e = getenv("CV_STENCIL_KCHUNK_SMALL");
    ks = e ? std::max(1, atoi(e)) : 2;
    e = getenv("CV_STENCIL_KCHUNK_MID");
    km = e ? std::max(1, atoi(e)) : 6;
    e = getenv("CV_STENCIL_KCHUNK_DEEP");
    kd = e ? std::max(1, atoi(e)) : 8;
  }
  return kc ? kc : (mz >= 192 ? kd : mz >= 64 ? km : ks);
}
// in a batch the small levels of many cases share one launch: 5 planes
// there measured best for the 64-case ensemble (102.7 ms per step against
// 103.0 with 4 and 104.1 with 8), and 8 on the middle tier (6 there cost
// 0.4 ms).  The tile shape does not change any point's arithmetic, so a
// case gets the same bits either way
static int stencil_kchunk_batch(int mz) {
  static int kb = -1, kbm = -1;
  if (kb < 0) {
    const char *e = getenv("CV_BATCH_KCHUNK_SMALL");
    kb = e ? std::max(1, atoi(e)) : 5;
    e = getenv("CV_BATCH_KCHUNK_MID");
    kbm = e ? std::max(1, atoi(e)) : 8;
  }
  if (mz >= 192 || getenv("CV_STENCIL_KCHUNK")) return stencil_kchunk(mz);
  return mz >= 64 ? kbm : kb;
}

// rows per stencil block (4 or 8; CV_STENCIL_ROWS64 / CV_STENCIL_ROWS32) and
// XCD-aware tile order (CV_STENCIL_REMAP); the defaults are the fastest
// measured at 512x256x256
static int env_int(const char *name, int dflt) {
  const char *e = getenv(name);
  return e ? atoi(e) : dflt;
}
template <class T>
static int stencil_rows() {
  static int ty = 0;
  if (!ty) ty = env_int(sizeof(T) == 8 ? "CV_STENCIL_ROWS64" : "CV_STENCIL_ROWS32", 4) == 8 ? 8 : 4;
  return ty;
}
static int stencil_pfb64() {
  static int v = -1;
  if (v < 0) v = env_int("CV_STENCIL_PFB64", 1) != 0;
  return v;
}
// tile order: 0 launch order, 1 one contiguous block range per XCD, R > 1
// interleaved runs of R tiles; default (-1 here) runs of two rows of x-tiles,
// the measured best at 512x256x256 (4-5 % over launch order)
static int stencil_remap(unsigned gx) {
  static int rm = -2;
  if (rm == -2) rm = env_int("CV_STENCIL_REMAP", -1);
  return rm >= 0 ? rm : (int)(2 * gx);
}

// algorithmic bytes per storage point of the stencil modes: x, 9 g, aj (+b)
// (+dinv), y in T, and the 4-byte operator word; hg: the 9 g, aj and the
// operator word as the 24-byte FP16 pack
template <class T>
static double stencil_bytes_pt(int mode, bool hg = false) {
  if (hg) return (double)(mode == 0 ? 2 : mode == 1 ? 3 : 4) * sizeof(T) + 24.0;
  return (double)(mode == 0 ? 12 : mode == 1 ? 13 : 14) * sizeof(T) + 4.0;
}

// the FP32 V-cycle reads the FP16 face coefficients
static bool half_coef(const cv_ctx *c) { return c->opt.precision == CV_MIXED16; }

// MODE 0: y = A x; 1: y = b - A x; 2: y = x + omega D^-1 (b - A x), on the
// cells of one (i+j+k) parity when color is 0 or 1 (the others keep x)
// tile rows of 32 lanes (31 cells per tile row, two rows per wavefront) by
// default, 4 rows per block: at 512x256x256 they waste fewer lanes on the
// storage width (514 = 16.6 x 31 against 8.2 x 63) and measured 3-4 %
// faster than rows of 64.  CV_STENCIL_TW=64 restores rows of 64 except on
// levels of at most CV_STENCIL_TW32_MAXMX columns; CV_STENCIL_TW32_ROWS: 4 or 8
static int stencil_tw(int mx) {
  static int all = -1, maxmx = -1;
  if (all < 0) {
    all = env_int("CV_STENCIL_TW", 32) == 32;
    maxmx = env_int("CV_STENCIL_TW32_MAXMX", 0);
  }
  return (all || mx <= maxmx) ? 32 : 64;
}
static int stencil_tw32_rows() {
  static int rows = 0;
  if (!rows) rows = env_int("CV_STENCIL_TW32_ROWS", 4) == 8 ? 8 : 4;
  return rows;
}

template <class T, int MODE, class TX = T>
static void stencil(cv_ctx *c, int l, const TX *x, const T *b, T *y, T omega = (T)0, int color = -1) {
  CvLevel &L = c->lev[l];
  OpV<T> A = opview<T>(c, l);
  const int kc = stencil_kchunk(L.d.mz);
  const int tw = stencil_tw(L.d.mx);
  const int ty = tw == 32 ? stencil_tw32_rows() : stencil_rows<T>();
  dim3 g((unsigned)((L.d.mx + tw - 2) / (tw - 1)), (unsigned)((L.d.my + ty - 1) / ty),
         (unsigned)((L.d.mz + kc - 1) / kc));
  const int remap = stencil_remap(g.x);
  const bool hg = sizeof(T) == 4 && half_coef(c) && L.gh;
  CvProfScope ps(c, MODE == 0 ? CV_PROF_APPLY : CV_PROF_RESID, l, sizeof(T) == 4 ? 1 : 0,
                 (double)L.d.n * (stencil_bytes_pt<T>(MODE, hg) - (double)(sizeof(T) - sizeof(TX))));
  const bool pfb = sizeof(T) == 4 || stencil_pfb64();
  if (tw == 32) {
    if constexpr (sizeof(T) == 4) {
      if (hg) {
        if (ty == 8)
          hipLaunchKernelGGL((k_stencil<T, MODE, 8, true, true, 32, TX>), g, dim3(32, 8), 0, c->stream, A, x, b, y, A.dinv,
                             omega, kc, remap, color);
        else
          hipLaunchKernelGGL((k_stencil<T, MODE, 4, true, true, 32, TX>), g, dim3(32, 4), 0, c->stream, A, x, b, y, A.dinv,
                             omega, kc, remap, color);
        return;
      }
    }
    if (ty == 8)
      hipLaunchKernelGGL((k_stencil<T, MODE, 8, true, false, 32, TX>), g, dim3(32, 8), 0, c->stream, A, x, b, y, A.dinv,
                         omega, kc, remap, color);
    else
      hipLaunchKernelGGL((k_stencil<T, MODE, 4, true, false, 32, TX>), g, dim3(32, 4), 0, c->stream, A, x, b, y, A.dinv,
                         omega, kc, remap, color);
    return;
  }
  if constexpr (sizeof(T) == 4) {
    if (hg) {
      if (ty == 4)
        hipLaunchKernelGGL((k_stencil<T, MODE, 4, true, true, 64, TX>), g, dim3(64, 4), 0, c->stream, A, x, b, y, A.dinv, omega,
                           kc, remap, color);
      else
        hipLaunchKernelGGL((k_stencil<T, MODE, 8, true, true, 64, TX>), g, dim3(64, 8), 0, c->stream, A, x, b, y, A.dinv, omega,
                           kc, remap, color);
      return;
    }
  }
  if (ty == 4 && pfb)
    hipLaunchKernelGGL((k_stencil<T, MODE, 4, true, false, 64, TX>), g, dim3(64, 4), 0, c->stream, A, x, b, y, A.dinv, omega, kc, remap,
                       color);
  else if (ty == 4)
    hipLaunchKernelGGL((k_stencil<T, MODE, 4, false, false, 64, TX>), g, dim3(64, 4), 0, c->stream, A, x, b, y, A.dinv, omega, kc, remap,
                       color);
  else if (pfb)
    hipLaunchKernelGGL((k_stencil<T, MODE, 8, true, false, 64, TX>), g, dim3(64, 8), 0, c->stream, A, x, b, y, A.dinv, omega, kc, remap,
                       color);
  else
    hipLaunchKernelGGL((k_stencil<T, MODE, 8, false, false, 64, TX>), g, dim3(64, 8), 0, c->stream, A, x, b, y, A.dinv, omega, kc, remap,
                       color);
}

// sum of part[0..nb) -> dsc[slot] (device)
static void final_sum(cv_ctx *c, const double *part, int nb, double *out, int nq) {
  hipLaunchKernelGGL(cv_k_final_sum, dim3(nq), dim3(BS), 0, c->stream, part, nb, out, nq);
}

int cv_levels_setup_nvert(cv_ctx *c, const double *nvert_host) {
  CvLevel &F = c->lev[c->nlev - 1];
  CV_CHECK(hipMemcpyAsync(F.nvert, nvert_host, F.d.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  for (int l = c->nlev - 1; l > 0; l--) {
    CvLevel &H = c->lev[l], &C = c->lev[l - 1];
    hipLaunchKernelGGL(k_nvrestrict1, dim3(nblk(C.d.n)), dim3(BS), 0, c->stream, idx_of(C), idx_of(H), C.d.isc,
                       C.d.jsc, C.d.ksc, (const double *)H.nvert, C.r, C.d.n);
    hipLaunchKernelGGL(k_nvrestrict2, dim3(nblk(C.d.n)), dim3(BS), 0, c->stream, idx_of(C), (const double *)C.r,
                       C.nvert, C.d.n);
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

__global__ static void k_count_fluid(const uint32_t *fl, long n, double *part) {
  __shared__ double sh[BS / 64];
  double s = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    if (fl[t] & 1u) s += 1.0;
  s = cv_block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------- FP16 face coefficients (CV_MIXED16) --------------
struct G9 { const double *p[9]; };
// max |g| over the 9 coefficient arrays: non-negative doubles order as
// their bit patterns, so an unsigned 64-bit max per wave does it
__global__ static void k_absmax9(G9 g, long n, unsigned long long *out) {
  double m = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
#pragma unroll
    for (int q = 0; q < 9; q++) m = cv_nmax(m, fabs(g.p[q][t]));
  for (int o = 32; o > 0; o >>= 1) m = cv_nmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}
__device__ inline uint32_t half_bits(double v) {
  return (uint32_t)__builtin_bit_cast(unsigned short, (_Float16)(float)v);
}
// the y-line factors of OpV::yh (a, ip * s) and OpV::yc (cp)
__global__ static void k_pack_line(const double *ya, const double *ip, const double *cp, double s, long n, uint32_t *yh,
                                   uint16_t *yc) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    yh[t] = half_bits(ya[t]) | (half_bits(ip[t] * s) << 16);
    yc[t] = (uint16_t)half_bits(cp[t]);
  }
}
__global__ static void k_absmax1(const double *a, long n, unsigned long long *out) {
  double m = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    m = cv_nmax(m, fabs(a[t]));
  for (int o = 32; o > 0; o >>= 1) m = cv_nmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// the two words per cell of OpV::hA / hB; the operator word has 16 bits
// (bit 0 and five per face axis)
__global__ static void k_pack_half(G9 g, const double *aj, const uint32_t *fl, double s, long n, uint4 *ha, uint2 *hb) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    auto h = [&](int q) { return half_bits(g.p[q][t] * s); };
    ha[t] = make_uint4(h(0) | (h(1) << 16), h(2) | (h(6) << 16), h(7) | (h(8) << 16),
                       __float_as_uint((float)(aj[t] / s)));
    hb[t] = make_uint2(h(3) | (h(4) << 16), h(5) | ((fl[t] & 0xffffu) << 16));
  }
}

// per level: 2^hexp puts max |g| in [2^13, 2^14) (FP16 holds up to 65504),
// g * 2^hexp in FP16 and aj * 2^-hexp in FP32, exact power-of-two scalings
static int half_ops(cv_ctx *c) {
  for (int l = 0; l < c->nlev; l++) {
    CvLevel &L = c->lev[l];
    const long n = L.d.n;
    int rc;
    void *p;
    if (!L.gh) {
      if ((rc = cv_alloc(c, &p, (size_t)32 * n))) return rc;
      L.gh = p;
    }
    G9 g;
    for (int q = 0; q < 9; q++) g.p[q] = L.g[q];
    CV_CHECK(hipMemsetAsync(c->part, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_absmax9, dim3((unsigned)std::min<long>(nblk(n), 1024)), dim3(BS), 0, c->stream, g, n,
                       (unsigned long long *)c->part);
    CV_CHECK(hipMemcpyAsync(c->red_host, c->part, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    const double gmax = c->red_host[0];
    if (!std::isfinite(gmax)) {
      cv_set_error("CV_MIXED16: level %d has a non-finite face coefficient", l);
      return CV_ERR_NONFINITE;
    }
    int e = 0;
    if (gmax > 0) frexp(gmax, &e);      // gmax < 2^e
    L.hexp = gmax > 0 ? 14 - e : 0;
    const double sc = ldexp(1.0, L.hexp);
    hipLaunchKernelGGL(k_pack_half, dim3(nblk(n)), dim3(BS), 0, c->stream, g, (const double *)L.aj,
                       (const uint32_t *)L.flags, sc, n, (uint4 *)L.gh, (uint2 *)((char *)L.gh + 16 * n));
    // line factors: 2^yexp puts max |ip| in [2^13, 2^14); |a|, |cp| < 1
    CV_CHECK(hipMemsetAsync(c->part, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_absmax1, dim3((unsigned)std::min<long>(nblk(n), 1024)), dim3(BS), 0, c->stream,
                       (const double *)L.yip, n, (unsigned long long *)c->part);
    CV_CHECK(hipMemcpyAsync(c->red_host, c->part, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    const double ipmax = c->red_host[0];
    if (!std::isfinite(ipmax)) {
      cv_set_error("CV_MIXED16: level %d has a non-finite y-line factor", l);
      return CV_ERR_NONFINITE;
    }
    int ey = 0;
    if (ipmax > 0) frexp(ipmax, &ey);
    L.yexp = ipmax > 0 ? 14 - ey : 0;
    hipLaunchKernelGGL(k_pack_line, dim3(nblk(n)), dim3(BS), 0, c->stream, (const double *)L.ylo, (const double *)L.yip,
                       (const double *)L.ycp, ldexp(1.0, L.yexp), n, (uint32_t *)((char *)L.gh + 24 * n),
                       (uint16_t *)((char *)L.gh + 28 * n));
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

int cv_poisson_prepare(cv_ctx *c) {
  for (int l = 0; l < c->nlev; l++) {
    CvLevel &L = c->lev[l];
    const long n = L.d.n;
    hipLaunchKernelGGL(k_flags, dim3(nblk(n)), dim3(BS), 0, c->stream, idx_of(L), L.d.xper, L.d.zper,
                       (const double *)L.nvert, L.flags, n);
    // line coefficients: lo -> ylo, diag -> r (scratch), up -> b (scratch)
    hipLaunchKernelGGL(k_linecoef, dim3(nblk(n)), dim3(BS), 0, c->stream, opview<double>(c, l), L.ylo, L.r, L.b, n);
    hipLaunchKernelGGL(k_linefactor, dim3(nblk((long)L.d.mx * L.d.mz)), dim3(BS), 0, c->stream, L.d.mx, L.d.my, L.d.mz,
                       L.ylo, (const double *)L.r, (const double *)L.b, L.yip, L.ycp, L.dinv);
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, L.ylof, (const double *)L.ylo, n);
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, L.yipf, (const double *)L.yip, n);
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, L.ycpf, (const double *)L.ycp, n);
    for (int q = 0; q < 9; q++)
      hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, L.gf[q], (const double *)L.g[q], n);
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, L.ajf, (const double *)L.aj, n);
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, L.dinvf, (const double *)L.dinv, n);
    hipLaunchKernelGGL(k_nvweight, dim3(nblk(n)), dim3(BS), 0, c->stream, (const double *)L.nvert, L.nwf, n);
    hipLaunchKernelGGL(k_count_fluid, dim3(CV_RED_GRID), dim3(BS), 0, c->stream, (const uint32_t *)L.flags, n, c->part);
    final_sum(c, c->part, CV_RED_GRID, c->dsc, 1);
    CV_CHECK(hipMemcpyAsync(c->red_host, c->dsc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    c->nfluid[l] = c->red_host[0];
  }
  CV_CHECK(hipGetLastError());
  if (half_coef(c)) {
    const int rc = half_ops(c);
    if (rc) c->opt.precision = CV_MIXED;     // no FP16 copies: the FP32 V-cycle
    return rc;
  }
  return CV_OK;
}

extern "C" int cv_set_nvert(cv_ctx *c, const double *nvert) {
  if (!c || !nvert) return CV_ERR_ARG;
  CV_CHECK(hipSetDevice(c->device));
  CV_CHECK(hipStreamSynchronize(c->stream));
  cv_ib_clear(c);                      // the IB face and cell lists follow nvert
  int rc = cv_levels_setup_nvert(c, nvert);
  if (rc) return rc;
  c->nvert_set = 1;
  if (c->geometry_set) return cv_poisson_prepare(c);
  return CV_OK;
}

// ---------------- nullspace projection on level l ------------------
template <class T>
static void nullspace_dev(cv_ctx *c, int l, T *x) {
  CvLevel &L = c->lev[l];
  const long n = L.d.n;
  const int nb = nred(n);
  hipLaunchKernelGGL(k_fluid_sum<T>, dim3(nb), dim3(BS), 0, c->stream, (const T *)x, (const uint32_t *)L.flags, n, c->part);
  final_sum(c, c->part, nb, c->dsc + 100, 1);
  double inv = c->nfluid[l] > 0 ? 1.0 / c->nfluid[l] : 0.0;
  hipLaunchKernelGGL(k_fluid_shift<T>, dim3(nblk(n)), dim3(BS), 0, c->stream, x, (const uint32_t *)L.flags,
                     (const double *)(c->dsc + 100), inv, n);
}

// ---------------- V-cycle ------------------------------------------
template <class T>
struct MgVecs {
  T *x[CV_MAXLEV], *b[CV_MAXLEV], *r[CV_MAXLEV];
};

template <class T>
static MgVecs<T> mgvecs(cv_ctx *c);
template <>
MgVecs<double> mgvecs<double>(cv_ctx *c) {
  MgVecs<double> v;
  for (int l = 0; l < c->nlev; l++) { v.x[l] = c->lev[l].x; v.b[l] = c->lev[l].b; v.r[l] = c->lev[l].r; }
  return v;
}
template <>
MgVecs<float> mgvecs<float>(cv_ctx *c) {
  MgVecs<float> v;
  for (int l = 0; l < c->nlev; l++) { v.x[l] = c->lev[l].xf; v.b[l] = c->lev[l].bf; v.r[l] = c->lev[l].rf; }
  return v;
}

// scratch for the forward elimination of the line solve, per precision
template <class T>
static T *line_tmp(cv_ctx *c, int l);
template <>
double *line_tmp<double>(cv_ctx *c, int l) { return c->lev[l].ltmp; }
template <>
float *line_tmp<float>(cv_ctx *c, int l) { return c->lev[l].ltmpf; }

// xo = xi + omega * T^{-1} r over every y-line of level l (T: the
// j-tridiagonal part of the operator); tmp: scratch of the long-line kernel
template <class T>
static void line_solve(cv_ctx *c, int l, const T *r, const T *xi, T *xo, T *tmp, T om, bool zg) {
  CvLevel &L = c->lev[l];
  OpV<T> A = opview<T>(c, l);
  const int rows = L.d.my - 2;
  const dim3 g((unsigned)(((long)L.d.mx * L.d.mz + 63) / 64));
  const dim3 gs((unsigned)((L.d.mx + 63) / 64), (unsigned)L.d.mz);   // k_yline_solve
  // bytes per point: r, the factors (FP16 packs: 6 bytes, else a, ip, cp), x in and out
  const bool hl = sizeof(T) == 4 && A.yh && rows <= 16 * 16;
  CvProfScope ps(c, CV_PROF_LINE, l, sizeof(T) == 4 ? 1 : 0, (double)L.d.n * (hl ? 18.0 : 6.0 * sizeof(T)));
  // FP64 keeps at most 16 rows per thread (32 would spill at 1024 threads)
  const int csmax = sizeof(T) == 8 ? 16 : 32;
  if (rows <= 16 * csmax) {
    const int cs = rows <= 16 * 8 ? 8 : rows <= 16 * 16 ? 16 : 32;
    const dim3 b(64, (unsigned)((rows + cs - 1) / cs));
    if constexpr (sizeof(T) == 4) {
      if (hl) {
        const T omh = om * A.ysc;
        // at most one block per CU: x loaded up front
        const bool px = g.x <= CV_PX_BLOCKS;
        if (cs == 8 && px)
          hipLaunchKernelGGL((k_yline_reg<T, 8, true, true>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo,
                             A.yip, A.ycp, xi, xo, omh, zg ? 1 : 0, A.yh, A.yc);
        else if (cs == 8)
          hipLaunchKernelGGL((k_yline_reg<T, 8, true>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo, A.yip,
                             A.ycp, xi, xo, omh, zg ? 1 : 0, A.yh, A.yc);
        else if (px)
          hipLaunchKernelGGL((k_yline_reg<T, 16, true, true>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo,
                             A.yip, A.ycp, xi, xo, omh, zg ? 1 : 0, A.yh, A.yc);
        else
          hipLaunchKernelGGL((k_yline_reg<T, 16, true>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo, A.yip,
                             A.ycp, xi, xo, omh, zg ? 1 : 0, A.yh, A.yc);
        return;
      }
    }
    if (cs == 8)
      hipLaunchKernelGGL((k_yline_reg<T, 8>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo, A.yip, A.ycp, xi,
                         xo, om, zg ? 1 : 0, (const uint32_t *)nullptr, (const uint16_t *)nullptr);
    else if (cs == 16)
      hipLaunchKernelGGL((k_yline_reg<T, 16>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo, A.yip, A.ycp, xi,
                         xo, om, zg ? 1 : 0, (const uint32_t *)nullptr, (const uint16_t *)nullptr);
    else
      hipLaunchKernelGGL((k_yline_reg<T, sizeof(T) == 8 ? 16 : 32>), g, b, 0, c->stream, L.d.mx, L.d.my, L.d.mz, r,
                         A.ylo, A.yip, A.ycp, xi, xo, om, zg ? 1 : 0, (const uint32_t *)nullptr, (const uint16_t *)nullptr);
  } else {
    hipLaunchKernelGGL(k_yline_solve<T>, gs, dim3(64, YL_NC), 0, c->stream, L.d.mx, L.d.my, L.d.mz, r, A.ylo, A.yip,
                       A.ycp, xi, xo, tmp, om, zg ? 1 : 0);
  }
}

// nsweep smoothing sweeps on level l; x in v.x[l] (zero_guess: x is 0), r
// used as ping-pong.  Red-black: each sweep is two damped half-sweeps, the
// cells of even (i+j+k) then the odd ones from the updated even ones; out of
// place, because the 19-point cross terms couple cells of one colour
template <class T>
static void smooth(cv_ctx *c, MgVecs<T> &v, int l, int nsweep, bool zero_guess) {
  CvLevel &L = c->lev[l];
  const long n = L.d.n;
  OpV<T> A = opview<T>(c, l);
  T om = (T)c->opt.omega;
  if (zero_guess && (nsweep == 0 || c->opt.smoother == CV_SMOOTH_RB))
    hipLaunchKernelGGL(k_zero<T>, dim3(nblk(n)), dim3(BS), 0, c->stream, v.x[l], n);
  for (int s = 0; s < nsweep; s++) {
    if (c->opt.smoother == CV_SMOOTH_RB) {
      stencil<T, 2>(c, l, (const T *)v.x[l], (const T *)v.b[l], v.r[l], om, 0);
      stencil<T, 2>(c, l, (const T *)v.r[l], (const T *)v.b[l], v.x[l], om, 1);
      continue;
    }
    if (c->opt.smoother == CV_SMOOTH_YLINE) {
      // residual into tmp, then the chunked line solve reads it and keeps
      // its forward sweep in the same array
      T *tmpv = line_tmp<T>(c, l);
      const bool zg = zero_guess && s == 0;
      const T *rr = (const T *)v.b[l];
      if (!zg) {
        stencil<T, 1>(c, l, (const T *)v.x[l], (const T *)v.b[l], tmpv);
        rr = tmpv;
      }
      line_solve<T>(c, l, rr, (const T *)v.x[l], v.r[l], tmpv, om, zg);
    } else {
      if (zero_guess && s == 0)
        hipLaunchKernelGGL(k_jacobi<T>, dim3(nblk(n)), dim3(BS), 0, c->stream, A, (const T *)v.x[l], (const T *)v.b[l],
                           v.r[l], om, 1, n);
      else
        stencil<T, 2>(c, l, (const T *)v.x[l], (const T *)v.b[l], v.r[l], om);
    }
    T *tmp = v.x[l]; v.x[l] = v.r[l]; v.r[l] = tmp;
  }
}

template <class T>
static void vcycle(cv_ctx *c, MgVecs<T> &v, int l) {
  CvLevel &L = c->lev[l];
  const long n = L.d.n;
  if (l == 0) {
    smooth<T>(c, v, 0, c->opt.coarse_sweeps, true);
    nullspace_dev<T>(c, 0, v.x[0]);
    return;
  }
  CvLevel &C = c->lev[l - 1];
  smooth<T>(c, v, l, c->opt.pre_sweeps, true);
  stencil<T, 1>(c, l, (const T *)v.x[l], (const T *)v.b[l], v.r[l]);
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(C.d.n)), dim3(BS), 0, c->stream, idx_of(C), idx_of(L), C.d.isc, C.d.jsc,
                     C.d.ksc, (const double *)L.nvert, (const float *)L.nwf, (const uint32_t *)C.flags, (const T *)v.r[l],
                     v.b[l - 1], C.d.n);
  vcycle<T>(c, v, l - 1);
  hipLaunchKernelGGL(k_prolong<T>, dim3(nblk(n)), dim3(BS), 0, c->stream, idx_of(L), idx_of(C), C.d.isc, C.d.jsc, C.d.ksc,
                     (const double *)C.nvert, (const uint32_t *)L.flags, (const T *)v.x[l - 1], v.x[l], 1, n);
  smooth<T>(c, v, l, c->opt.post_sweeps, false);
}

// z = M^{-1} v on the finest level (FP64 or FP32 V-cycle), then the nullspace rule
// Z = M^-1 vin with the nullspace rule.  Mixed precisions: the FP32 V-cycle
// writes its result straight into zout, which then holds the FP32
// preconditioned vector (the FP64 apply and the solution update widen it
// exactly), and the nullspace rule runs on it in FP32 with FP64 sums
// b_ready: the FP32 right-hand side of the mixed V-cycle already holds vin
static void precondition(cv_ctx *c, const double *vin, void *zout, bool b_ready = false) {
  const int top = c->nlev - 1;
  CvLevel &F = c->lev[top];
  const long n = F.d.n;
  if (c->opt.precision != CV_FP64) {
    float *const zf = (float *)zout;
    MgVecs<float> v = mgvecs<float>(c);
    float *const own_x = v.x[top], *const own_r = v.r[top];
    if (!b_ready) hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BS), 0, c->stream, v.b[top], vin, n);
    v.x[top] = zf;
    vcycle<float>(c, v, top);
    if (v.x[top] != zf)
      hipLaunchKernelGGL(k_copy_t<float>, dim3(nblk(n)), dim3(BS), 0, c->stream, zf, (const float *)v.x[top], n);
    v.x[top] = own_x;
    v.r[top] = own_r;
    for (int l = 0; l < c->nlev; l++) { c->lev[l].xf = v.x[l]; c->lev[l].rf = v.r[l]; }
    nullspace_dev<float>(c, top, zf);
    return;
  } else {
    // the finest level reads its right-hand side from vin and starts its
    // x/r ping-pong in zout: after an even number of sweeps the result is
    // in zout already (the V-cycle only reads b, and the arithmetic does
    // not depend on which buffer holds x)
    double *const zd = (double *)zout;
    MgVecs<double> v = mgvecs<double>(c);
    double *const own_x = v.x[top], *const own_r = v.r[top];
    v.b[top] = const_cast<double *>(vin);
    v.x[top] = zd;
    vcycle<double>(c, v, top);
    if (v.x[top] != zd)
      hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(BS), 0, c->stream, zd, (const double *)v.x[top], n);
    v.x[top] = own_x;
    v.r[top] = own_r;
    for (int l = 0; l < c->nlev; l++) { c->lev[l].x = v.x[l]; c->lev[l].r = v.r[l]; }
  }
  nullspace_dev<double>(c, top, (double *)zout);
}

static int ensure_krylov(cv_ctx *c) {
  int m = c->opt.restart;
  if (m < 1) m = 1;
  if (m > MAXV - 2) m = MAXV - 2;
  c->opt.restart = m;
  if (c->krylov_m >= m) return CV_OK;
  const long n = c->lev[c->nlev - 1].d.n;
  void *p;
  int rc;
  if (c->V) { cv_free(c, c->V); c->V = nullptr; }
  if (c->Z) { cv_free(c, c->Z); c->Z = nullptr; }
  if ((rc = cv_alloc(c, &p, (size_t)(m + 1) * n * sizeof(double)))) return rc;
  c->V = (double *)p;
  if ((rc = cv_alloc(c, &p, (size_t)m * n * sizeof(double)))) return rc;
  c->Z = (double *)p;
  c->krylov_m = m;
  return CV_OK;
}

// one classical Gram-Schmidt pass of w against V_0..V_{nv-1}: h and ||w||^2
// to dsc[0..nv], then w -= V h and ||w||^2 to dsc[36]
static void gram_schmidt_pass(cv_ctx *c, double *w, const double *V, int nv, long n) {
  double *hd = c->dsc;
  {
    CvProfScope ps(c, CV_PROF_ORTH, c->nlev - 1, 0, (double)n * 8.0 * (nv + 1));
    hipLaunchKernelGGL(mdot_tbl[nv - 1], dim3(nred(n)), dim3(BS), 0, c->stream, (const double *)w, V, n, n, c->part);
  }
  final_sum(c, c->part, nred(n), hd, nv + 1);
  hipLaunchKernelGGL(orth_tbl[nv - 1], dim3(nred(n)), dim3(BS), 0, c->stream, w, V, n, (const double *)hd, n,
                     c->part);
  final_sum(c, c->part, nred(n), hd + 36, 1);
}

// FGMRES(m) for A phi = b on the finest level; phi starts at 0
static int fgmres(cv_ctx *c, const double *b, double *x, int *its_out, double *rnorm_out, double *bnorm_out) {
  const int top = c->nlev - 1;
  CvLevel &F = c->lev[top];
  const long n = F.d.n;
  int rc = ensure_krylov(c);
  if (rc) return rc;
  const int m = c->opt.restart;
  double *V = c->V, *Z = c->Z, *w = c->w;
  // the preconditioned vectors of the mixed solve are the FP32 V-cycle's
  // output, kept in FP32 (packed in the first half of Z's allocation)
  const bool zf32 = c->opt.precision != CV_FP64;
  double *hd = c->dsc;              // device: h[0..33]
  double *hh = c->red_host;         // pinned host
  // ||b||
  hipLaunchKernelGGL(k_sumsq, dim3(nred(n)), dim3(BS), 0, c->stream, b, n, c->part);
  final_sum(c, c->part, nred(n), hd + 40, 1);
  CV_CHECK(hipMemcpyAsync(hh, hd + 40, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  const double bnorm = std::sqrt(hh[0]);
  *bnorm_out = bnorm;
  if (!std::isfinite(bnorm)) {
    cv_set_error("FGMRES: the right-hand side is not finite");
    return CV_ERR_NONFINITE;
  }
  const double tol = std::fmax(c->opt.rtol * bnorm, c->opt.atol);
  hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n)), dim3(BS), 0, c->stream, x, n);
  int its = 0;
  double rnorm = bnorm;
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  // r = b (x = 0)
  hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(BS), 0, c->stream, w, b, n);
  double beta = bnorm;
  while (beta > tol && its < c->opt.max_it) {
    hh[0] = 1.0 / beta;
    CV_CHECK(hipMemcpyAsync(hd + 48, hh, sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (zf32)
      hipLaunchKernelGGL(k_scale_to_f32, dim3(nblk(n)), dim3(BS), 0, c->stream, V, (const double *)w,
                         (const double *)(hd + 48), c->lev[top].bf, n);
    else
      hipLaunchKernelGGL(k_scale_to, dim3(nblk(n)), dim3(BS), 0, c->stream, V, (const double *)w,
                         (const double *)(hd + 48), n);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int jj = 0;
    for (; jj < m && its < c->opt.max_it; jj++) {
      double *vj = V + (size_t)jj * n;
      if (zf32) {
        // V_jj's FP32 copy was written with V_jj itself (k_scale_to_f32)
        float *zj = (float *)Z + (size_t)jj * n;
        precondition(c, vj, zj, true);
        stencil<double, 0, float>(c, top, (const float *)zj, (const double *)nullptr, w);
      } else {
        double *zj = Z + (size_t)jj * n;
        precondition(c, vj, zj);
        stencil<double, 0>(c, top, (const double *)zj, (const double *)nullptr, w);
      }
      const int nv = jj + 1;
      // first CGS pass
      gram_schmidt_pass(c, w, V, nv, n);
      CV_CHECK(hipMemcpyAsync(hh, hd, 40 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      CV_CHECK(hipStreamSynchronize(c->stream));
      double wn2 = hh[nv], hn2 = hh[36];
      for (int q = 0; q < nv; q++) H[q * m + jj] = hh[q];
      if (hn2 < 1e-2 * wn2) {      // DGKS: second classical pass when ||w|| fell below 0.1 of itself
        gram_schmidt_pass(c, w, V, nv, n);
        CV_CHECK(hipMemcpyAsync(hh, hd, 40 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CV_CHECK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < nv; q++) H[q * m + jj] += hh[q];
        hn2 = hh[36];
      }
      double hn = std::sqrt(std::fmax(hn2, 0.0));
      H[(jj + 1) * m + jj] = hn;
      if (!std::isfinite(hn)) return CV_ERR_NONFINITE;
      if (hn > 0) {
        hh[0] = 1.0 / hn;
        CV_CHECK(hipMemcpyAsync(hd + 48, hh, sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (zf32)
          hipLaunchKernelGGL(k_scale_to_f32, dim3(nblk(n)), dim3(BS), 0, c->stream, V + (size_t)(jj + 1) * n,
                             (const double *)w, (const double *)(hd + 48), c->lev[top].bf, n);
        else
          hipLaunchKernelGGL(k_scale_to, dim3(nblk(n)), dim3(BS), 0, c->stream, V + (size_t)(jj + 1) * n,
                             (const double *)w, (const double *)(hd + 48), n);
      }
      for (int q = 0; q < jj; q++) {
        double a = H[q * m + jj], bb = H[(q + 1) * m + jj];
        H[q * m + jj] = cs[q] * a + sn[q] * bb;
        H[(q + 1) * m + jj] = -sn[q] * a + cs[q] * bb;
      }
      double a = H[jj * m + jj], bb = H[(jj + 1) * m + jj];
      double den = std::sqrt(a * a + bb * bb);
      cs[jj] = den == 0 ? 1.0 : a / den;
      sn[jj] = den == 0 ? 0.0 : bb / den;
      H[jj * m + jj] = cs[jj] * a + sn[jj] * bb;
      H[(jj + 1) * m + jj] = 0.0;
      g[jj + 1] = -sn[jj] * g[jj];
      g[jj] = cs[jj] * g[jj];
      its++;
      rnorm = std::fabs(g[jj + 1]);
      if (rnorm <= tol || hn == 0) { jj++; break; }
    }
    for (int q = jj - 1; q >= 0; q--) {
      double s = g[q];
      for (int t = q + 1; t < jj; t++) s -= H[q * m + t] * y[t];
      y[q] = s / H[q * m + q];
    }
    for (int q = 0; q < jj; q++) hh[q] = y[q];
    CV_CHECK(hipMemcpyAsync(hd, hh, jj * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (zf32)
      hipLaunchKernelGGL(k_xupdate<float>, dim3(nblk(n)), dim3(BS), 0, c->stream, x, (const float *)Z, (long)n,
                         (const double *)hd, jj, n);
    else
      hipLaunchKernelGGL(k_xupdate<double>, dim3(nblk(n)), dim3(BS), 0, c->stream, x, (const double *)Z, (long)n,
                         (const double *)hd, jj, n);
    // true residual
    stencil<double, 1>(c, top, (const double *)x, b, w);
    hipLaunchKernelGGL(k_sumsq, dim3(nred(n)), dim3(BS), 0, c->stream, (const double *)w, n, c->part);
    final_sum(c, c->part, nred(n), hd + 40, 1);
    CV_CHECK(hipMemcpyAsync(hh, hd + 40, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    beta = std::sqrt(hh[0]);
    rnorm = beta;
    if (!std::isfinite(beta)) return CV_ERR_NONFINITE;
  }
  CV_CHECK(hipGetLastError());
  *its_out = its;
  *rnorm_out = rnorm;
  return rnorm <= tol ? CV_OK : CV_ERR_NOT_CONVERGED;
}

// ====================================================================
// PoissonRHS (poisson.c:1509-1576)
// ====================================================================
__device__ __forceinline__ static void poisson_rhs_body(Idx I, const double *ux, const double *uy, const double *uz, const double *aj,
                                     const double *nv, double *b, double dt, double st, long n, double *part, int bid, int nb) {
  __shared__ double sh[BS / 64];
  double s = 0.0;
  for (long t = (long)bid * BS + threadIdx.x; t < n; t += (long)nb * BS) {
    int i, j, k;
    I.dec(t, i, j, k);
    double r;
    if (i == 0 || i == I.mx - 1 || j == 0 || j == I.my - 1 || k == 0 || k == I.mz - 1) r = 0.;
    else if (nv[t] > 0.1) r = 0;
    else
      r = -(ux[t] - ux[t - 1] + uy[t] - uy[t - I.mx] + uz[t] - uz[t - (long)I.mx * I.my]) / dt * aj[t] / st * 1.5;
    b[t] = r;
    s += r / aj[t] * dt / 1.5;
  }
  s = cv_block_sum(s, sh);
  if (threadIdx.x == 0) part[bid] = s;
}
__global__ static void k_poisson_rhs(Idx I, const double *ux, const double *uy, const double *uz, const double *aj,
                                     const double *nv, double *b, double dt, double st, long n, double *part) { poisson_rhs_body(I, ux, uy, uz, aj, nv, b, dt, st, n, part, blockIdx.x, gridDim.x); }


int cv_poisson_rhs_dev(cv_ctx *c, double *rhs_sum) {
  CvLevel &F = c->lev[c->nlev - 1];
  const int nb = nred(F.d.n);
  hipLaunchKernelGGL(k_poisson_rhs, dim3(nb), dim3(BS), 0, c->stream, idx_of(F), (const double *)c->ucont[0],
                     (const double *)c->ucont[1], (const double *)c->ucont[2], (const double *)F.aj,
                     (const double *)F.nvert, c->bvec, c->dt, c->st, F.d.n, c->part);
  final_sum(c, c->part, nb, c->dsc + 104, 1);
  if (rhs_sum) {
    CV_CHECK(hipMemcpyAsync(c->red_host + 100, c->dsc + 104, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    *rhs_sum = c->red_host[100];
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// ====================================================================
// VolumeFlux_rev (poisson.c:3712-3944)
// ====================================================================
__device__ static inline double vnorm(const double *x, const double *y, const double *z, long t) {
  return sqrt(x[t] * x[t] + y[t] * y[t] + z[t] * z[t]);
}

// mode 0: sum flux/area into part[2*b], mode 1: apply correction dsc[slot]
__device__ __forceinline__ static void ibm_flux_body(Idx I, const double *nv, double *ux, double *uy, double *uz, const double *cx,
                                  const double *cy, const double *cz, const double *ex, const double *ey, const double *ez,
                                  const double *zx, const double *zy, const double *zz, int mode, const double *corr_p,
                                  long n, double *part, int bid, int nb) {
  __shared__ double sh[BS / 64];
  const double ibmval = 1.5;
  const double corr = mode ? corr_p[0] : 0.0;
  double f = 0.0, a = 0.0;
  const long sj = I.mx, sk = (long)I.mx * I.my;
  // one point as the reference's loop body visits it
  auto visit = [&](long t) {
    int i, j, k;
    I.dec(t, i, j, k);
    if (!(i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1 && k > 0 && k < I.mz - 1)) return;
    const double n0 = nv[t];
    const long nb[3] = {t + 1, t + sj, t + sk};
    const bool lim[3] = {i < I.mx - 2, j < I.my - 2, k < I.mz - 2};
    double *u[3] = {ux, uy, uz};
    const double *vx[3] = {cx, ex, zx}, *vy[3] = {cy, ey, zy}, *vz[3] = {cz, ez, zz};
    for (int d = 0; d < 3; d++) {
      const double nn = nv[nb[d]];
      if (n0 < 0.1 && nn > ibmval - 0.4 && nn < ibmval && lim[d]) {
        double ar = vnorm(vx[d], vy[d], vz[d], t);
        if (mode) u[d][t] -= ar * corr; else { f += u[d][t]; a += ar; }
      }
    }
    if (n0 > ibmval - 0.4 && n0 < ibmval) {
      for (int d = 0; d < 3; d++) {
        if (nv[nb[d]] < 0.1 && lim[d]) {
          double ar = vnorm(vx[d], vy[d], vz[d], t);
          if (mode) u[d][t] += ar * corr; else { f -= u[d][t]; a += ar; }
        }
      }
    }
  };
  auto band = [ibmval](double v) { return v > ibmval - 0.4 && v < ibmval; };
  // the thread's points t, t + S, ... in order, four at a time: their
  // nvert loads are issued together and only the points beside the body
  // (a superset of those visit changes anything for) are visited, in the
  // same order as one at a time, so the sums keep their bits
  const long S = (long)nb * BS;
  constexpr int U = 4;
  for (long t0 = (long)bid * BS + threadIdx.x; t0 < n; t0 += U * S) {
    unsigned cand = 0u;
#pragma unroll
    for (int q = 0; q < U; q++) {
      const long t = t0 + q * S;
      if (t < n) {
        int i, j, k;
        I.dec(t, i, j, k);
        if (i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1 && k > 0 && k < I.mz - 1) {
          const double n0 = nv[t], a1 = nv[t + 1], a2 = nv[t + sj], a3 = nv[t + sk];
          if ((n0 < 0.1 && (band(a1) || band(a2) || band(a3))) || (band(n0) && (a1 < 0.1 || a2 < 0.1 || a3 < 0.1)))
            cand |= 1u << q;
        }
      }
    }
#pragma unroll 1
    for (int q = 0; q < U; q++)
      if ((cand >> q) & 1u) visit(t0 + q * S);
  }
  if (mode) return;
  double sf = cv_block_sum(f, sh);
  if (threadIdx.x == 0) part[bid] = sf;
  double sa = cv_block_sum(a, sh);
  if (threadIdx.x == 0) part[nb + bid] = sa;
}
__global__ static void k_ibm_flux(Idx I, const double *nv, double *ux, double *uy, double *uz, const double *cx,
                                  const double *cy, const double *cz, const double *ex, const double *ey, const double *ez,
                                  const double *zx, const double *zy, const double *zz, int mode, const double *corr_p,
                                  long n, double *part) { ibm_flux_body(I, nv, ux, uy, uz, cx, cy, cz, ex, ey, ez, zx, zy, zz, mode, corr_p, n, part, blockIdx.x, gridDim.x); }


__device__ __forceinline__ static void ibm_corr_body(const double *fa, double *corr) {
  corr[0] = fa[1] > 1.e-15 ? (fa[0] + 0.0) / fa[1] : 0.0;   // FluxIntpSum = 0
}
__global__ static void k_ibm_corr(const double *fa, double *corr) { ibm_corr_body(fa, corr); }

int cv_volumeflux_rev_dev(cv_ctx *c, double *fa4) {
  CvLevel &F = c->lev[c->nlev - 1];
  CvMetrics &M = c->m;
  const int nb = nred(F.d.n);
  for (int pass = 0; pass < 3; pass++) {
    int mode = pass == 1 ? 1 : 0;
    hipLaunchKernelGGL(k_ibm_flux, dim3(nb), dim3(BS), 0, c->stream, idx_of(F), (const double *)F.nvert,
                       c->ucont[0], c->ucont[1], c->ucont[2], (const double *)M.csi[0], (const double *)M.csi[1],
                       (const double *)M.csi[2], (const double *)M.eta[0], (const double *)M.eta[1],
                       (const double *)M.eta[2], (const double *)M.zet[0], (const double *)M.zet[1],
                       (const double *)M.zet[2], mode, (const double *)(c->dsc + 108), F.d.n, c->part);
    if (!mode) {
      final_sum(c, c->part, nb, c->dsc + (pass == 0 ? 112 : 114), 2);
      if (pass == 0) hipLaunchKernelGGL(k_ibm_corr, dim3(1), dim3(1), 0, c->stream, (const double *)(c->dsc + 112), c->dsc + 108);
    }
  }
  if (fa4) {
    CV_CHECK(hipMemcpyAsync(c->red_host + 104, c->dsc + 112, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CV_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 4; q++) fa4[q] = c->red_host[104 + q];
  }
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// ====================================================================
// PoissonSolver_MG (poisson.c:1775-2017)
// ====================================================================
int cv_poisson_solve_dev(cv_ctx *c, cv_step_stats *st) {
  if (!c->geometry_set || !c->nvert_set) { cv_set_error("geometry/nvert not set"); return CV_ERR_STATE; }
  double fa[4];
  int rc = cv_volumeflux_rev_dev(c, fa);
  if (rc) return rc;
  double rs = 0;
  rc = cv_poisson_rhs_dev(c, &rs);
  if (rc) return rc;
  int its = 0;
  double rn = 0, bn = 0;
  rc = fgmres(c, c->bvec, c->phi, &its, &rn, &bn);
  if (rc && rc != CV_ERR_NOT_CONVERGED) return rc;
  nullspace_dev<double>(c, c->nlev - 1, c->phi);
  CV_CHECK(hipGetLastError());
  if (st) {
    st->ksp_its = its; st->ksp_rnorm = rn; st->ksp_bnorm = bn; st->rhs_sum = rs;
    st->ibm_flux0 = fa[0]; st->ibm_area0 = fa[1]; st->ibm_flux1 = fa[2]; st->ibm_area1 = fa[3];
    st->converged = rc == CV_OK;
  }
  return rc;
}

// ====================================================================
// UpdatePressure (poisson.c:2908-3248)
// ====================================================================
__device__ __forceinline__ static void p_add_body(Idx I, double *p, const double *phi, long n, long t0_, long st_) {
  for (long t = t0_; t < n; t += st_) {
    int i, j, k;
    I.dec(t, i, j, k);
    if (i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1 && k > 0 && k < I.mz - 1) p[t] += phi[t];
  }
}
__global__ static void k_p_add(Idx I, double *p, const double *phi, long n) { p_add_body(I, p, phi, n, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x); }

// x planes: i=0 <- mx-2, i=mx-1 <- 1 over (j,k) in [j0,j1)x[k0,k1)
__device__ __forceinline__ static void pplane_x_body(Idx I, double *a, double *b, int j0, int j1, int k0, int k1, long t0_, long st_) {
  long nj = j1 - j0, nk = k1 - k0;
  for (long t = t0_; t < nj * nk; t += st_) {
    int j = j0 + (int)(t % nj), k = k0 + (int)(t / nj);
    long p0 = I.at(0, j, k), pm = I.at(I.mx - 2, j, k), q0 = I.at(I.mx - 1, j, k), q1 = I.at(1, j, k);
    a[p0] = a[pm]; a[q0] = a[q1];
    if (b) { b[p0] = b[pm]; b[q0] = b[q1]; }
  }
}
__global__ static void k_pplane_x(Idx I, double *a, double *b, int j0, int j1, int k0, int k1) { pplane_x_body(I, a, b, j0, j1, k0, k1, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x); }

__device__ __forceinline__ static void pplane_z_body(Idx I, double *a, double *b, int i0, int i1, int j0, int j1, long t0_, long st_) {
  long ni = i1 - i0, nj = j1 - j0;
  for (long t = t0_; t < ni * nj; t += st_) {
    int i = i0 + (int)(t % ni), j = j0 + (int)(t / ni);
    long p0 = I.at(i, j, 0), pm = I.at(i, j, I.mz - 2), q0 = I.at(i, j, I.mz - 1), q1 = I.at(i, j, 1);
    a[p0] = a[pm]; a[q0] = a[q1];
    if (b) { b[p0] = b[pm]; b[q0] = b[q1]; }
  }
}
__global__ static void k_pplane_z(Idx I, double *a, double *b, int i0, int i1, int j0, int j1) { pplane_z_body(I, a, b, i0, i1, j0, j1, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x); }


int cv_update_pressure_dev(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  Idx I = idx_of(F);
  hipLaunchKernelGGL(k_p_add, dim3(nblk(F.d.n)), dim3(BS), 0, c->stream, I, c->p, (const double *)c->phi, F.d.n);
  const int mx = F.d.mx, my = F.d.my, mz = F.d.mz;
  if (F.d.xper)
    hipLaunchKernelGGL(k_pplane_x, dim3(nblk((long)my * mz)), dim3(BS), 0, c->stream, I, c->p, c->phi, 1, my - 1, 1, mz - 1);
  if (F.d.zper)
    hipLaunchKernelGGL(k_pplane_z, dim3(nblk((long)mx * my)), dim3(BS), 0, c->stream, I, c->p, c->phi, 1, mx - 1, 1, my - 1);
  if (F.d.xper)
    hipLaunchKernelGGL(k_pplane_x, dim3(nblk((long)my * mz)), dim3(BS), 0, c->stream, I, c->p, c->phi, 0, my, 0, mz);
  if (F.d.zper)
    hipLaunchKernelGGL(k_pplane_z, dim3(nblk((long)mx * my)), dim3(BS), 0, c->stream, I, c->p, c->phi, 0, mx, 0, my);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// ====================================================================
// Projection (poisson.c:2281-2906), face corrections only; the caller
// follows with Contra2Cart and GhostNodeVelocity
// ====================================================================
struct FaceMet {
  const double *g[3];   // (xi_d . xi_t) / J of one face family d, t = 0..2 (the operator's g[3d + t])
};

// the tangential difference of p along axis t on the face of axis d of the
// cell at storage index c (strides of the axes: 1, mx, mx*my)
__device__ static inline double tang_p(const double *p, Idx I, int d, int t, long c, int mode) {
  const long so[3] = {1, (long)I.mx, (long)I.mx * I.my};
  const long o = so[d], f = so[t];
  if (mode == XM_LOWER) return (p[c] + p[c + o] - p[c - f] - p[c + o - f]) * 0.5;
  if (mode == XM_UPPER) return (p[c + f] + p[c + o + f] - p[c] - p[c + o]) * 0.5;
  if (mode == XM_CENTRAL) return (p[c + f] + p[c + o + f] - p[c - f] - p[c + o - f]) * 0.25;
  return 0.0;
}

__device__ static inline void project_face(Idx I, int d, int i, int j, int k, const double *phi, const double *nv,
                                           const uint32_t *fl, FaceMet M, double *u, double dt, double st) {
  const long t = I.at(i, j, k);
  const long up = t + (d == 0 ? 1 : d == 1 ? (long)I.mx : (long)I.mx * I.my);
  if (nv[t] + nv[up] != 0.0) return;
  uint32_t w = fl[t];
  double dp[3];
  int q = 0;
  for (int tt = 0; tt < 3; tt++) {
    if (tt == d) { dp[tt] = phi[up] - phi[t]; continue; }
    int mode = (w >> (2 + 5 * d + 2 * q)) & 3u; q++;
    dp[tt] = tang_p(phi, I, d, tt, t, mode);
  }
  // the coefficients of poisson.c:2420-2431 (dot products of the face
  // metrics times the face 1/J) are the operator's g of this face
  const double g0 = M.g[0][t], g1 = M.g[1][t], g2 = M.g[2][t];
  u[t] -= (dp[0] * g0 + dp[1] * g1 + dp[2] * g2) * dt * st / 1.5;
}

// tang_p from loaded values: p0 = p[c], po = p[c+o], pf and pof = p[c+f],
// p[c+o+f], pmf and pomf = p[c-f], p[c+o-f]; the same terms in the same
// order as tang_p, so the same bits
__device__ static inline double tang_v(int mode, double p0, double po, double pf, double pof, double pmf,
                                       double pomf) {
  const bool lo = mode == XM_LOWER, up = mode == XM_UPPER;
  const double a = lo ? p0 : pf, b = lo ? po : pof;
  const double c = up ? p0 : pmf, d = up ? po : pomf;
  const double v = (a + b - c - d) * (lo || up ? 0.5 : 0.25);
  return mode == XM_DROP ? 0.0 : v;
}

// the 16 points of phi the three faces' rules read around an interior cell
// (all inside the grid) and the nvert of the cell and its +i, +j, +k
// neighbours
struct ProjPts {
  double p0, pe, pn, pt, pw, ps, pb, pen, pes, pet, peb, pnw, pnt, pnb, ptw, pts;
  double n0, ne, nn, nt;
};

// the three faces of an interior cell t (1 <= i, j, k <= m-2) as three
// project_face calls (fx, fy, fz: which of them to update), from points
// loaded before the fluid and mode tests so that the loads overlap
__device__ static inline void project_faces(long t, bool fx, bool fy, bool fz, const ProjPts &q, uint32_t w,
                                            const FaceMet &Mi, const FaceMet &Mj, const FaceMet &Mk, double *ux,
                                            double *uy, double *uz, double dt, double st) {
  auto md = [w](int d, int k) { return (int)((w >> (2 + 5 * d + 2 * k)) & 3u); };
  if (fx && q.n0 + q.ne == 0.0) {
    const double dp0 = q.pe - q.p0;
    const double dp1 = tang_v(md(0, 0), q.p0, q.pe, q.pn, q.pen, q.ps, q.pes);
    const double dp2 = tang_v(md(0, 1), q.p0, q.pe, q.pt, q.pet, q.pb, q.peb);
    const double g0 = Mi.g[0][t], g1 = Mi.g[1][t], g2 = Mi.g[2][t];
    ux[t] -= (dp0 * g0 + dp1 * g1 + dp2 * g2) * dt * st / 1.5;
  }
  if (fy && q.n0 + q.nn == 0.0) {
    const double dp0 = tang_v(md(1, 0), q.p0, q.pn, q.pe, q.pen, q.pw, q.pnw);
    const double dp1 = q.pn - q.p0;
    const double dp2 = tang_v(md(1, 1), q.p0, q.pn, q.pt, q.pnt, q.pb, q.pnb);
    const double g0 = Mj.g[0][t], g1 = Mj.g[1][t], g2 = Mj.g[2][t];
    uy[t] -= (dp0 * g0 + dp1 * g1 + dp2 * g2) * dt * st / 1.5;
  }
  if (fz && q.n0 + q.nt == 0.0) {
    const double dp0 = tang_v(md(2, 0), q.p0, q.pt, q.pe, q.pet, q.pw, q.ptw);
    const double dp1 = tang_v(md(2, 1), q.p0, q.pt, q.pn, q.pnt, q.ps, q.pts);
    const double dp2 = q.pt - q.p0;
    const double g0 = Mk.g[0][t], g1 = Mk.g[1][t], g2 = Mk.g[2][t];
    uz[t] -= (dp0 * g0 + dp1 * g1 + dp2 * g2) * dt * st / 1.5;
  }
}

// the points of cell t that the cell below (t - sk) did not load: its plane
// k+1 points are this cell's own plane, its own plane this cell's k-1
__device__ static inline void proj_load_top(const double *phi, const double *nv, long t, long sj, long sk, ProjPts &q) {
  q.pt = phi[t + sk]; q.pnt = phi[t + sj + sk]; q.pet = phi[t + 1 + sk]; q.ptw = phi[t + sk - 1];
  q.pts = phi[t + sk - sj];
  q.pen = phi[t + 1 + sj]; q.pes = phi[t + 1 - sj]; q.pnw = phi[t + sj - 1];
  q.ne = nv[t + 1]; q.nn = nv[t + sj]; q.nt = nv[t + sk];
}

__device__ static inline void project_cell(Idx I, long t, bool fx, bool fy, bool fz, const double *phi,
                                           const double *nv, const uint32_t *fl, const FaceMet &Mi,
                                           const FaceMet &Mj, const FaceMet &Mk, double *ux, double *uy, double *uz,
                                           double dt, double st) {
  const long sj = (long)I.mx, sk = (long)I.mx * I.my;
  ProjPts q;
  q.p0 = phi[t]; q.pe = phi[t + 1]; q.pn = phi[t + sj]; q.pw = phi[t - 1]; q.ps = phi[t - sj];
  q.pb = phi[t - sk]; q.pnb = phi[t + sj - sk]; q.peb = phi[t + 1 - sk];
  q.n0 = nv[t];
  proj_load_top(phi, nv, t, sj, sk, q);
  project_faces(t, fx, fy, fz, q, fl[t], Mi, Mj, Mk, ux, uy, uz, dt, st);
}

__device__ __forceinline__ static void projection_body(Idx I, int xper, int zper, const double *phi, const double *nv, const uint32_t *fl,
                                    FaceMet Mi, FaceMet Mj, FaceMet Mk, double *ux, double *uy, double *uz, double dt,
                                    double st, long n, long t0_, long st_) {
  const int i_end = xper ? I.mx - 1 : I.mx - 2, j_end = I.my - 2, k_end = zper ? I.mz - 1 : I.mz - 2;
  IdxWalk w(I, t0_, st_);
  for (long t = t0_; t < n; t += st_, w.next()) {
    const int i = w.i, j = w.j, k = w.k;
    bool jk = j > 0 && j < I.my - 1 && k > 0 && k < I.mz - 1, ij = i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1;
    if (jk && i > 0 && i < I.mx - 1)
      project_cell(I, t, i < i_end, j < j_end, k < k_end, phi, nv, fl, Mi, Mj, Mk, ux, uy, uz, dt, st);
    if (xper && i == 0 && jk) project_face(I, 0, 0, j, k, phi, nv, fl, Mi, ux, dt, st);
    if (zper && k == 0 && ij) project_face(I, 2, i, j, 0, phi, nv, fl, Mk, uz, dt, st);
  }
}


// The projection as a k-march (64 x PJ_TY threads walk PJ_KC planes): the
// same faces as projection_body, each cell taking 8 of its phi points and
// its nvert from the cell below it in registers
#define PJ_TY 4
#define PJ_KC 16
__global__ static void __launch_bounds__(64 * PJ_TY)
    k_projection_march(Idx I, int xper, int zper, const double *phi, const double *nv, const uint32_t *fl, FaceMet Mi,
                       FaceMet Mj, FaceMet Mk, double *ux, double *uy, double *uz, double dt, double st) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * PJ_TY + threadIdx.y;
  if (i >= I.mx || j >= I.my) return;
  const int k0 = blockIdx.z * PJ_KC, k1 = min(k0 + PJ_KC, I.mz);
  const int i_end = xper ? I.mx - 1 : I.mx - 2, j_end = I.my - 2, k_end = zper ? I.mz - 1 : I.mz - 2;
  const bool iin = i > 0 && i < I.mx - 1, jin = j > 0 && j < I.my - 1;
  const long sj = (long)I.mx, sk = (long)I.mx * I.my;
  bool have = false;
  ProjPts c{};   // the cell below's points
  for (int k = k0; k < k1; k++) {
    const bool kin = k > 0 && k < I.mz - 1;
    if (xper && i == 0 && jin && kin) project_face(I, 0, 0, j, k, phi, nv, fl, Mi, ux, dt, st);
    if (zper && k == 0 && iin && jin) project_face(I, 2, i, j, 0, phi, nv, fl, Mk, uz, dt, st);
    if (!(iin && jin && kin)) { have = false; continue; }
    const long t = I.at(i, j, k);
    ProjPts q;
    if (have) {
      q.p0 = c.pt; q.pe = c.pet; q.pn = c.pnt; q.pw = c.ptw; q.ps = c.pts;
      q.pb = c.p0; q.pnb = c.pn; q.peb = c.pe;
      q.n0 = c.nt;
    } else {
      q.p0 = phi[t]; q.pe = phi[t + 1]; q.pn = phi[t + sj]; q.pw = phi[t - 1]; q.ps = phi[t - sj];
      q.pb = phi[t - sk]; q.pnb = phi[t + sj - sk]; q.peb = phi[t + 1 - sk];
      q.n0 = nv[t];
    }
    proj_load_top(phi, nv, t, sj, sk, q);
    project_faces(t, i < i_end, j < j_end, k < k_end, q, fl[t], Mi, Mj, Mk, ux, uy, uz, dt, st);
    c = q;
    have = true;
  }
}

int cv_projection_dev(cv_ctx *c) {
  CvLevel &F = c->lev[c->nlev - 1];
  FaceMet Mi{{F.g[0], F.g[1], F.g[2]}};
  FaceMet Mj{{F.g[3], F.g[4], F.g[5]}};
  FaceMet Mk{{F.g[6], F.g[7], F.g[8]}};
  const dim3 grid((unsigned)((F.d.mx + 63) / 64), (unsigned)((F.d.my + PJ_TY - 1) / PJ_TY),
                  (unsigned)((F.d.mz + PJ_KC - 1) / PJ_KC));
  hipLaunchKernelGGL(k_projection_march, grid, dim3(64, PJ_TY), 0, c->stream, idx_of(F), F.d.xper, F.d.zper,
                     (const double *)c->phi, (const double *)F.nvert, (const uint32_t *)F.flags, Mi, Mj, Mk,
                     c->ucont[0], c->ucont[1], c->ucont[2], c->dt, c->st);
  CV_CHECK(hipGetLastError());
  int rc = cv_contra2cart_dev(c);
  if (rc) return rc;
  return cv_ghost_update(c, 0);
}

// ====================================================================
// Divergence (main.c:469-593): max and first index of the maximum
// ====================================================================
// (v2, i2) beats (v1, i1): larger, or equal with the smaller index; a NaN
// beats every number, so a field holding a NaN reports NaN (and the first
// NaN's index) instead of hiding it
__host__ __device__ inline bool div_beats(double v2, long i2, double v1, long i1) {
  const bool n2 = v2 != v2, n1 = v1 != v1;
  if (n2 != n1) return n2;
  if (!n2 && v2 != v1) return v2 > v1;
  return i2 < i1;
}
// block bid of nb: the largest |div| of its grid-stride share and its first
// index (ties go to the smaller index, so the result does not depend on
// the partition)
// S: element stride of the flux components (1: planar arrays, 3: the
// interleaved Cmpnts of the DMDA layout, ux = u, uy = u + 1, uz = u + 2)
template <int S>
__device__ __forceinline__ void divmax_block_s(Idx I, const double *ux, const double *uy, const double *uz,
                                               const double *aj, const double *nv, long n, double *pv, double *pi,
                                               int bid, int nb) {
  __shared__ double sv[BS];
  __shared__ long si[BS];
  double best = -1.0;
  long bi = 0;
  const long sj = I.mx, sk = (long)I.mx * I.my;
  for (long t = bid * (long)BS + threadIdx.x; t < n; t += (long)nb * BS) {
    int i, j, k;
    I.dec(t, i, j, k);
    double dv = 0.0;
    if (i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1 && k > 0 && k < I.mz - 1) {
      dv = fabs((ux[S * t] - ux[S * (t - 1)] + uy[S * t] - uy[S * (t - sj)] + uz[S * t] - uz[S * (t - sk)]) * aj[t]);
      if (nv[t] + nv[t + sk] + nv[t - sk] + nv[t + sj] + nv[t - sj] + nv[t + 1] + nv[t - 1] > 0.1) dv = 0.;
    }
    if (div_beats(dv, t, best, bi)) { best = dv; bi = t; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double v2 = sv[threadIdx.x + s];
      long i2 = si[threadIdx.x + s];
      if (div_beats(v2, i2, sv[threadIdx.x], si[threadIdx.x])) { sv[threadIdx.x] = v2; si[threadIdx.x] = i2; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { pv[bid] = sv[0]; pi[bid] = (double)si[0]; }
}
__device__ __forceinline__ void divmax_block(Idx I, const double *ux, const double *uy, const double *uz,
                                             const double *aj, const double *nv, long n, double *pv, double *pi,
                                             int bid, int nb) {
  divmax_block_s<1>(I, ux, uy, uz, aj, nv, n, pv, pi, bid, nb);
}

// divmax_block_s<1> as k-marches: block b takes tiles b, b + nb, ... of 64 x
// DV_TY columns by DV_KC planes, and each column's cell takes uz and two
// nvert planes from the cell below it.  Every point is visited once, and
// the maximum with its first index does not depend on the visiting order,
// so the partials reduce to the same (value, index) as the flat pass
#define DV_TY 4
#define DV_KC 16
__global__ static void __launch_bounds__(64 * DV_TY)
    k_divmax_march(Idx I, const double *ux, const double *uy, const double *uz, const double *aj, const double *nv,
                   double *pv, double *pi) {
  static_assert(64 * DV_TY == BS, "one reduction slot per thread");
  __shared__ double sv[BS];
  __shared__ long si[BS];
  const int tx = (I.mx + 63) / 64, ty = (I.my + DV_TY - 1) / DV_TY, tz = (I.mz + DV_KC - 1) / DV_KC;
  const long ntile = (long)tx * ty * tz;
  const long sj = I.mx, sk = (long)I.mx * I.my;
  double best = -1.0;
  long bi = 0;
  for (long tb = blockIdx.x; tb < ntile; tb += gridDim.x) {
    const int bx = (int)(tb % tx), by = (int)((tb / tx) % ty), bz = (int)(tb / ((long)tx * ty));
    const int i = bx * 64 + threadIdx.x, j = by * DV_TY + threadIdx.y;
    if (i >= I.mx || j >= I.my) continue;
    const int k0 = bz * DV_KC, k1 = min(k0 + DV_KC, I.mz);
    const bool ij = i > 0 && i < I.mx - 1 && j > 0 && j < I.my - 1;
    bool have = false;
    double c_uz = 0., c_n0 = 0., c_nt = 0.;   // the cell below's uz[t], nv[t], nv[t + sk]
    for (int k = k0; k < k1; k++) {
      const long t = I.at(i, j, k);
      double dv = 0.0;
      if (ij && k > 0 && k < I.mz - 1) {
        const double uzb = have ? c_uz : uz[t - sk], nb_ = have ? c_n0 : nv[t - sk], n0 = have ? c_nt : nv[t];
        const double uzt = uz[t], nt = nv[t + sk];
        dv = fabs((ux[t] - ux[t - 1] + uy[t] - uy[t - sj] + uzt - uzb) * aj[t]);
        if (n0 + nt + nb_ + nv[t + sj] + nv[t - sj] + nv[t + 1] + nv[t - 1] > 0.1) dv = 0.;
        c_uz = uzt; c_n0 = n0; c_nt = nt;
        have = true;
      } else {
        have = false;
      }
      if (div_beats(dv, t, best, bi)) { best = dv; bi = t; }
    }
  }
  const int tid = threadIdx.y * 64 + threadIdx.x;
  sv[tid] = best;
  si[tid] = bi;
  __syncthreads();
  for (int s2 = BS / 2; s2 > 0; s2 >>= 1) {
    if (tid < s2) {
      double v2 = sv[tid + s2];
      long i2 = si[tid + s2];
      if (div_beats(v2, i2, sv[tid], si[tid])) { sv[tid] = v2; si[tid] = i2; }
    }
    __syncthreads();
  }
  if (tid == 0) { pv[blockIdx.x] = sv[0]; pi[blockIdx.x] = (double)si[0]; }
}

int cv_divergence_dev(cv_ctx *c, double *maxdiv, long *idx) {
  CvLevel &F = c->lev[c->nlev - 1];
  const int nb = CV_RED_GRID;
  hipLaunchKernelGGL(k_divmax_march, dim3(nb), dim3(64, DV_TY), 0, c->stream, idx_of(F), (const double *)c->ucont[0],
                     (const double *)c->ucont[1], (const double *)c->ucont[2], (const double *)F.aj,
                     (const double *)F.nvert, c->part, c->part + nb);
  std::vector<double> hv(2 * nb);
  CV_CHECK(hipMemcpyAsync(hv.data(), c->part, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  double best = -1.0;
  long bi = 0;
  for (int b = 0; b < nb; b++) {
    long ii = (long)hv[nb + b];
    if (div_beats(hv[b], ii, best, bi)) { best = hv[b]; bi = ii; }
  }
  if (maxdiv) *maxdiv = best;
  if (idx) *idx = bi;
  return CV_OK;
}

// ====================================================================
// Divergence without device allocation (cv_divergence_noalloc): the block
// partials live in a file-scope device array, one copy per GPU, that every
// call on that GPU shares
// ====================================================================
#define CV_NOALLOC_PARTS CV_RED_GRID       // (value, index) pairs: 16 KiB per GPU
#define CV_NOALLOC_DEVICES 64
static __device__ __align__(256) double scratch[2 * CV_NOALLOC_PARTS];
static double *g_scratch[CV_NOALLOC_DEVICES];   // device address of each GPU's copy
static std::mutex g_scratch_mu;

// the address of this GPU's copy of scratch, looked up on the first call
// on the device (which is current); in host code the array's name is only
// a stand-in for it
static int noalloc_scratch(int device, double **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      cv_set_error("cv_divergence_noalloc: hipGetSymbolAddress(scratch) -> %s", hipGetErrorString(e));
      return CV_ERR_HIP;
    }
    g_scratch[device] = (double *)p;
  }
  *out = g_scratch[device];
  return CV_OK;
}

// block partials of Divergence over the DMDA layout (Cmpnts interleaved)
__global__ static void k_divmax_dmda(Idx I, const double *u, const double *aj, const double *nv, long n,
                                     double *pv, double *pi) {
  divmax_block_s<3>(I, u, u + 1, u + 2, aj, nv, n, pv, pi, blockIdx.x, gridDim.x);
}
// the largest partial and its first index (the order cv_divergence_dev
// applies on the host) into out[0], out[1]
__global__ static void k_divmax_final(const double *pv, const double *pi, int nb, double *out) {
  __shared__ double sv[BS];
  __shared__ long si[BS];
  double best = -1.0;
  long bi = 0;
  for (int b = threadIdx.x; b < nb; b += BS) {
    const long ii = (long)pi[b];
    if (div_beats(pv[b], ii, best, bi)) { best = pv[b]; bi = ii; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s && div_beats(sv[threadIdx.x + s], si[threadIdx.x + s], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + s];
      si[threadIdx.x] = si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = sv[0]; out[1] = (double)si[0]; }
}

extern "C" int cv_divergence_noalloc(int device, const double *ucont, const double *aj, const double *nvert, int mx,
                                     int my, int mz, void *stream, double *out) {
  if (!ucont || !aj || !nvert || !out || device < 0 || device >= CV_NOALLOC_DEVICES) {
    cv_set_error("cv_divergence_noalloc: null pointer or device outside 0..%d", CV_NOALLOC_DEVICES - 1);
    return CV_ERR_ARG;
  }
  if (mx < 3 || my < 3 || mz < 3 || (long)mx * my * mz >= (1L << 31)) {
    cv_set_error("cv_divergence_noalloc: storage %d x %d x %d outside 3^3 .. 2^31 points", mx, my, mz);
    return CV_ERR_ARG;
  }
  CV_CHECK(hipSetDevice(device));
  double *part = nullptr;
  int rc = noalloc_scratch(device, &part);
  if (rc) return rc;
  const long n = (long)mx * my * mz;
  hipStream_t s = (hipStream_t)stream;
  // every block writes its own pair, so nothing needs clearing first
  hipLaunchKernelGGL(k_divmax_dmda, dim3(CV_NOALLOC_PARTS), dim3(BS), 0, s, Idx{mx, my, mz}, ucont, aj, nvert, n,
                     part, part + CV_NOALLOC_PARTS);
  hipLaunchKernelGGL(k_divmax_final, dim3(1), dim3(BS), 0, s, (const double *)part,
                     (const double *)(part + CV_NOALLOC_PARTS), CV_NOALLOC_PARTS, out);
  CV_CHECK(hipGetLastError());
  return CV_OK;
}

// ====================================================================
// C ABI
// ====================================================================
static int need_ready(cv_ctx *c) {
  if (!c) return CV_ERR_ARG;
  if (!c->geometry_set || !c->nvert_set) { cv_set_error("cv: geometry and nvert must be set first"); return CV_ERR_STATE; }
  CV_CHECK(hipSetDevice(c->device));
  return CV_OK;
}

extern "C" int cv_set_solver(cv_ctx *c, const cv_solver_opts *o) {
  if (!c || !o) return CV_ERR_ARG;
  if (o->restart < 1 || o->restart > MAXV - 2 || o->max_it < 1 || o->rtol < 0 || o->pre_sweeps < 0 ||
      o->post_sweeps < 0 || o->coarse_sweeps < 1 ||
      (o->precision != CV_FP64 && o->precision != CV_MIXED && o->precision != CV_MIXED16) ||
      o->omega <= 0 || (o->smoother != CV_SMOOTH_JACOBI && o->smoother != CV_SMOOTH_RB && o->smoother != CV_SMOOTH_YLINE)) {
    cv_set_error("cv_set_solver: invalid options");
    return CV_ERR_ARG;
  }
  const cv_solver_opts prev = c->opt;
  c->opt = *o;
  if (half_coef(c) && c->geometry_set && c->nvert_set) {
    CV_CHECK(hipSetDevice(c->device));
    const int rc = half_ops(c);
    if (rc) c->opt = prev;
    return rc;
  }
  return CV_OK;
}

extern "C" int cv_poisson_rhs(cv_ctx *c, double *b_out) {
  int rc = need_ready(c);
  if (rc) return rc;
  double rs;
  if ((rc = cv_poisson_rhs_dev(c, &rs))) return rc;
  c->last.rhs_sum = rs;
  if (b_out) return cv_get_field(c, CV_F_B, -1, b_out);
  return CV_OK;
}

extern "C" int cv_poisson_apply(cv_ctx *c, int level, const double *x, double *y) {
  int rc = need_ready(c);
  if (rc) return rc;
  if (level < 0) level = c->nlev - 1;
  if (level >= c->nlev || !x || !y) return CV_ERR_ARG;
  CvLevel &L = c->lev[level];
  CV_CHECK(hipMemcpyAsync(L.b, x, L.d.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  stencil<double, 0>(c, level, (const double *)L.b, (const double *)nullptr, L.r);
  CV_CHECK(hipGetLastError());
  CV_CHECK(hipMemcpyAsync(y, L.r, L.d.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  return CV_OK;
}

extern "C" int cv_poisson_solve(cv_ctx *c, cv_step_stats *st) {
  int rc = need_ready(c);
  if (rc) return rc;
  cv_step_stats s{};
  rc = cv_poisson_solve_dev(c, &s);
  c->last.ksp_its = s.ksp_its; c->last.ksp_rnorm = s.ksp_rnorm; c->last.ksp_bnorm = s.ksp_bnorm;
  c->last.rhs_sum = s.rhs_sum; c->last.converged = s.converged;
  if (st) *st = s;
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cv_volumeflux_rev(cv_ctx *c, double *fa4) {
  int rc = need_ready(c);
  if (rc) return rc;
  double tmp[4];
  rc = cv_volumeflux_rev_dev(c, fa4 ? fa4 : tmp);
  return rc;
}

extern "C" int cv_update_pressure(cv_ctx *c) {
  int rc = need_ready(c);
  if (rc) return rc;
  rc = cv_update_pressure_dev(c);
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cv_projection(cv_ctx *c) {
  int rc = need_ready(c);
  if (rc) return rc;
  rc = cv_projection_dev(c);
  CV_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cv_divergence(cv_ctx *c, double *maxdiv, long *idx) {
  int rc = need_ready(c);
  if (rc) return rc;
  return cv_divergence_dev(c, maxdiv, idx);
}

extern "C" int cv_mg_restrict(cv_ctx *c, int lc, const double *xf, double *xc) {
  int rc = need_ready(c);
  if (rc) return rc;
  if (lc < 0 || lc >= c->nlev - 1 || !xf || !xc) return CV_ERR_ARG;
  CvLevel &C = c->lev[lc], &F = c->lev[lc + 1];
  CV_CHECK(hipMemcpyAsync(F.r, xf, F.d.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_restrict<double>, dim3(nblk(C.d.n)), dim3(BS), 0, c->stream, idx_of(C), idx_of(F), C.d.isc,
                     C.d.jsc, C.d.ksc, (const double *)F.nvert, (const float *)F.nwf, (const uint32_t *)C.flags,
                     (const double *)F.r, C.b, C.d.n);
  CV_CHECK(hipMemcpyAsync(xc, C.b, C.d.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  return CV_OK;
}

extern "C" int cv_mg_prolong(cv_ctx *c, int lf, const double *xc, double *xf) {
  int rc = need_ready(c);
  if (rc) return rc;
  if (lf < 1 || lf >= c->nlev || !xf || !xc) return CV_ERR_ARG;
  CvLevel &C = c->lev[lf - 1], &F = c->lev[lf];
  CV_CHECK(hipMemcpyAsync(C.r, xc, C.d.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_prolong<double>, dim3(nblk(F.d.n)), dim3(BS), 0, c->stream, idx_of(F), idx_of(C), C.d.isc,
                     C.d.jsc, C.d.ksc, (const double *)C.nvert, (const uint32_t *)F.flags, (const double *)C.r, F.b, 0,
                     F.d.n);
  CV_CHECK(hipMemcpyAsync(xf, F.b, F.d.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  return CV_OK;
}

extern "C" int cv_nullspace_project(cv_ctx *c, int level, double *x) {
  int rc = need_ready(c);
  if (rc) return rc;
  if (level < 0) level = c->nlev - 1;
  if (level >= c->nlev || !x) return CV_ERR_ARG;
  CvLevel &L = c->lev[level];
  CV_CHECK(hipMemcpyAsync(L.b, x, L.d.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  nullspace_dev<double>(c, level, L.b);
  CV_CHECK(hipMemcpyAsync(x, L.b, L.d.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CV_CHECK(hipStreamSynchronize(c->stream));
  return CV_OK;
}

#include "cv_batch_solve.h"
