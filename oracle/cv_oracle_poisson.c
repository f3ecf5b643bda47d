/*
 * cv_oracle_poisson.c -- TEST INFRASTRUCTURE (CPU oracle): pressure Poisson
 * path of the fractional step.
 *
 *   o_poisson_lhs      PoissonLHSNew      poisson.c:600-1507  (assembled CSR)
 *   o_poisson_rhs      PoissonRHS         poisson.c:1509-1576
 *   o_nullspace        MatNullSpaceRemove(has_cnst) + PoissonNullSpaceFunction
 *                                          poisson.c:372-563, 1950-1954
 *   fully_blocked      FullyBlocked       poisson.c:1578-1660
 *   o_restrict         MyRestriction      poisson.c:129-226
 *   o_interp           MyInterpolation    poisson.c:228-370
 *   o_poisson_solve    PoissonSolver_MG   poisson.c:1775-2017: FGMRES right-
 *                      preconditioned by a multiplicative V-cycle whose level
 *                      smoothers are Richardson + block-Jacobi ILU(0) (one
 *                      block per thread, as PCBJACOBI gives one per rank) and
 *                      whose coarse solve is one block-Jacobi ILU(0) sweep.
 *   o_update_pressure  UpdatePressure     poisson.c:2908-3248
 *   o_projection       Projection         poisson.c:2281-2906
 *   o_volumeflux_rev   VolumeFlux_rev     poisson.c:3712-3944
 *   o_divergence       Divergence         main.c:469-593
 *
 * Parity pins: the operator against the textbook 7-point Laplacian on
 * uniform grids (cross coefficients vanish) with its Neumann eigenpairs,
 * A*1 = 0 on fluid rows of a warped grid with a body, R.1 = P.1 = 1, the
 * nullspace rule in closed form, the solve against a dense least-squares
 * solution, a residual-scale divergence after a step
 * (tests/test_oracle_pins.py); PoissonRHS of a known divergence, Projection
 * subtracting (face area).grad phi for a linear phi on uniform and sheared
 * lattices, UpdatePressure p += phi, VolumeFlux_rev cancelling the net
 * fluid-to-IB flux with one correction per unit face area, Divergence on
 * planted maxima and ties (tests/test_oracle_poisson.py).
 * tests/test_oracle_mutations.py plants a mistake in each of these routines
 * and checks that a pin fails.  The converged Phi is normalised by the
 * reference's nullspace rule (zero mean over interior fluid cells, zero on
 * boundary/solid cells).
 */
#include "cv_oracle_int.h"
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* cross-derivative mode of one tangential direction of a face (the rule of
 * poisson.c:800-837 and its eleven siblings; identical in Projection and
 * FormFunction1).  hi/lo_pair: nvert sums of the two cells on the upper /
 * lower tangential side; at_hi/at_lo: t==m-2 / t==1; per: axis periodic. */
enum { XM_CENTRAL = 0, XM_LOWER = 1, XM_UPPER = 2, XM_DROP = 3 };
int o_xmode(double hi_pair, double lo_pair, int at_hi, int at_lo, int per) {
  if ((at_hi && !per) || hi_pair > 0.1) {
    if (lo_pair < 0.1 && (!at_lo || per)) return XM_LOWER;
    return XM_DROP;
  } else if ((at_lo && !per) || lo_pair > 0.1) {
    if (hi_pair < 0.1) return XM_UPPER;
    return XM_DROP;
  }
  return XM_CENTRAL;
}

/* ------------------------------------------------------------------ */
/* PoissonLHSNew (poisson.c:600-1507)                                   */
/* ------------------------------------------------------------------ */
/* 19-point stencil offsets (di,dj,dk) in the reference's CP..BW order
 * (poisson.c:577-595) */
static const int ST[19][3] = {
  {0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
  {1,1,0},{1,-1,0},{-1,1,0},{-1,-1,0},{0,1,1},{0,1,-1},{0,-1,1},{0,-1,-1},
  {1,0,1},{1,0,-1},{-1,0,1},{-1,0,-1}};

static int st_index(int di, int dj, int dk) {
  for (int m = 0; m < 19; m++)
    if (ST[m][0] == di && ST[m][1] == dj && ST[m][2] == dk) return m;
  return -1;
}

static double gface(const ocmp *a, const ocmp *b, double ajf) {
  return (a->x * b->x + a->y * b->y + a->z * b->z) * ajf;
}

/* Adds the flux of one face to vol[] with sign sg (+1 for the upper face of
 * the cell, -1 for the lower face).  d: face normal axis; fi,fj,fk: face
 * index (cell on its lower side); g[3]: the face's g_d1, g_d2, g_d3. */
static void add_face(const olevel *L, double vol[19], int d, int sg, int i, int j, int k,
                     const double g[3]) {
  const ofield *nv = &L->nvert;
  /* the two cells sharing the face: (i,j,k) and its neighbour at sg along d */
  int o[3] = {0, 0, 0}; o[d] = sg;
  int lo_is_nb = (sg < 0);  /* for the lower face the neighbour is the lower cell */
  int m_ax[3] = {L->mx, L->my, L->mz};
  int ijk[3] = {i, j, k};
  for (int t = 0; t < 3; t++) {
    if (t == d) {
      /* normal term: sg * g_dd * (x_nb - x_c) * ... written as the reference:
       * upper face: CP -= g, NB += g; lower face: CP -= g, NB += g */
      int mnb = st_index(o[0], o[1], o[2]);
      vol[0] -= g[t];
      vol[mnb] += g[t];
      continue;
    }
    /* tangential axis t */
    int up[3] = {0, 0, 0}, dn[3] = {0, 0, 0}; up[t] = 1; dn[t] = -1;
    double hi = S(nv, k + up[2], j + up[1], i + up[0]) + S(nv, k + up[2] + o[2], j + up[1] + o[1], i + up[0] + o[0]);
    double lo = S(nv, k + dn[2], j + dn[1], i + dn[0]) + S(nv, k + dn[2] + o[2], j + dn[1] + o[1], i + dn[0] + o[0]);
    int at_hi = (ijk[t] == m_ax[t] - 2), at_lo = (ijk[t] == 1);
    int per = (t == 0) ? (L->bctype[0] == 7) : (t == 1) ? (L->bctype[2] == 7) : (L->bctype[4] == 7);
    int mode = o_xmode(hi, lo, at_hi, at_lo, per);
    (void)lo_is_nb;
    /* stencil positions: cell row c (offset 0) and neighbour row (offset o),
     * at tangential offsets +1 / 0 / -1 */
    int pc_up = st_index(up[0], up[1], up[2]), pn_up = st_index(o[0] + up[0], o[1] + up[1], o[2] + up[2]);
    int pc_dn = st_index(dn[0], dn[1], dn[2]), pn_dn = st_index(o[0] + dn[0], o[1] + dn[1], o[2] + dn[2]);
    int pc = 0, pn = st_index(o[0], o[1], o[2]);
    double gg = g[t];
    double s = (double)sg;
    if (mode == XM_LOWER) {
      vol[pc] += s * gg * 0.5; vol[pn] += s * gg * 0.5;
      vol[pc_dn] -= s * gg * 0.5; vol[pn_dn] -= s * gg * 0.5;
    } else if (mode == XM_UPPER) {
      vol[pc_up] += s * gg * 0.5; vol[pn_up] += s * gg * 0.5;
      vol[pc] -= s * gg * 0.5; vol[pn] -= s * gg * 0.5;
    } else if (mode == XM_CENTRAL) {
      vol[pc_up] += s * gg * 0.25; vol[pn_up] += s * gg * 0.25;
      vol[pc_dn] -= s * gg * 0.25; vol[pn_dn] -= s * gg * 0.25;
    }
  }
}

/* Face coefficient triplets (poisson.c:699-735) */
static void gcoef(const olevel *L, int d, int i, int j, int k, double g[3]) {
  if (d == 0) {
    const ocmp *cs = &V(&L->icsi, k, j, i), *et = &V(&L->ieta, k, j, i), *ze = &V(&L->izet, k, j, i);
    double a = S(&L->iaj, k, j, i);
    g[0] = gface(cs, cs, a); g[1] = gface(et, cs, a); g[2] = gface(ze, cs, a);
  } else if (d == 1) {
    const ocmp *cs = &V(&L->jcsi, k, j, i), *et = &V(&L->jeta, k, j, i), *ze = &V(&L->jzet, k, j, i);
    double a = S(&L->jaj, k, j, i);
    g[0] = gface(cs, et, a); g[1] = gface(et, et, a); g[2] = gface(ze, et, a);
  } else {
    const ocmp *cs = &V(&L->kcsi, k, j, i), *et = &V(&L->keta, k, j, i), *ze = &V(&L->kzet, k, j, i);
    double a = S(&L->kaj, k, j, i);
    g[0] = gface(cs, ze, a); g[1] = gface(et, ze, a); g[2] = gface(ze, ze, a);
  }
}

static int wrapi(int i, int m, int per) {
  if (!per) return i;
  if (i == m - 1) return 1;      /* poisson.c:1331-1334 */
  if (i == 0) return m - 2;
  return i;
}

void o_poisson_lhs(octx *c, int level) {
  olevel *L = &c->lev[level];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const int xper = L->bctype[0] == 7, zper = L->bctype[4] == 7;
  const int x_end = xper ? mx - 1 : mx - 2, x_str = xper ? 0 : 1;
  const int y_end = my - 2, y_str = 1;
  const int z_end = zper ? mz - 1 : mz - 2, z_str = zper ? 0 : 1;
  const double threshold = 0.1;
  const long N = (long)mx * my * mz;
  free(L->rowptr); free(L->col); free(L->val);
  L->nrows = (int)N;
  L->rowptr = (int *)malloc((N + 1) * sizeof(int));
  L->col = (int *)malloc(N * 19 * sizeof(int));
  L->val = (double *)malloc(N * 19 * sizeof(double));
  /* rows are independent: count first, then fill in parallel */
  int *cnt = (int *)malloc(N * sizeof(int));
  const ofield *nv = &L->nvert;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < N; r++) {
    int i = (int)(r % mx), j = (int)((r / mx) % my), k = (int)(r / ((long)mx * my));
    int fluid = !(i == 0 || i == mx - 1 || j == 0 || j == my - 1 || k == 0 || k == mz - 1) &&
                !(S(nv, k, j, i) > 0.1);
    cnt[r] = fluid ? 19 : 1;
  }
  L->rowptr[0] = 0;
  for (long r = 0; r < N; r++) L->rowptr[r + 1] = L->rowptr[r] + cnt[r];
  free(cnt);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < N; r++) {
    int i = (int)(r % mx), j = (int)((r / mx) % my), k = (int)(r / ((long)mx * my));
    int base = L->rowptr[r];
    if (L->rowptr[r + 1] - base == 1) {     /* boundary or solid row: identity */
      L->col[base] = (int)r; L->val[base] = 1.0;
      continue;
    }
    double vol[19];
    for (int m = 0; m < 19; m++) vol[m] = 0.;
    double g[3];
    if (S(nv, k, j, i + 1) < threshold && i != x_end) { gcoef(L, 0, i, j, k, g); add_face(L, vol, 0, +1, i, j, k, g); }
    if (S(nv, k, j, i - 1) < threshold && i != x_str) { gcoef(L, 0, i - 1, j, k, g); add_face(L, vol, 0, -1, i, j, k, g); }
    if (S(nv, k, j + 1, i) < threshold && j != y_end) { gcoef(L, 1, i, j, k, g); add_face(L, vol, 1, +1, i, j, k, g); }
    if (S(nv, k, j - 1, i) < threshold && j != y_str) { gcoef(L, 1, i, j - 1, k, g); add_face(L, vol, 1, -1, i, j, k, g); }
    if (S(nv, k + 1, j, i) < threshold && k != z_end) { gcoef(L, 2, i, j, k, g); add_face(L, vol, 2, +1, i, j, k, g); }
    if (S(nv, k - 1, j, i) < threshold && k != z_str) { gcoef(L, 2, i, j, k - 1, g); add_face(L, vol, 2, -1, i, j, k, g); }
    double aj = S(&L->aj, k, j, i);
    for (int m = 0; m < 19; m++) vol[m] *= -aj;       /* poisson.c:1325-1327 */
    /* column indices with the periodic wrap of poisson.c:1329-1450, sorted */
    int cols[19];
    for (int m = 0; m < 19; m++) {
      int ii = wrapi(i + ST[m][0], mx, xper), jj = j + ST[m][1], kk = wrapi(k + ST[m][2], mz, zper);
      cols[m] = (int)(ii + (long)mx * (jj + (long)my * kk));
    }
    int ord[19];
    for (int m = 0; m < 19; m++) ord[m] = m;
    for (int a = 1; a < 19; a++) {
      int t = ord[a], b = a - 1;
      while (b >= 0 && cols[ord[b]] > cols[t]) { ord[b + 1] = ord[b]; b--; }
      ord[b + 1] = t;
    }
    for (int m = 0; m < 19; m++) { L->col[base + m] = cols[ord[m]]; L->val[base + m] = vol[ord[m]]; }
  }
}

void o_csr_matvec(const olevel *L, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (long r = 0; r < L->nrows; r++) {
    double s = 0.;
    for (int q = L->rowptr[r]; q < L->rowptr[r + 1]; q++) s += L->val[q] * x[L->col[q]];
    y[r] = s;
  }
}

/* ------------------------------------------------------------------ */
/* PoissonRHS (poisson.c:1509-1576)                                     */
/* ------------------------------------------------------------------ */
void o_poisson_rhs(octx *c, ofield *B) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  ofield *uc = &c->ucont, *nv = &L->nvert, *aj = &L->aj;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        double r;
        if (i == 0 || i == mx - 1 || j == 0 || j == my - 1 || k == 0 || k == mz - 1) r = 0.;
        else if (S(nv, k, j, i) > 0.1) r = 0.;
        else
          r = -(V(uc,k,j,i).x - V(uc,k,j,i-1).x + V(uc,k,j,i).y - V(uc,k,j-1,i).y +
                V(uc,k,j,i).z - V(uc,k-1,j,i).z) / c->dt * S(aj, k, j, i) / c->st * COEF_TIME_ACCURACY;
        S(B, k, j, i) = r;
      }
  double sum = 0.;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++)
        sum += S(B, k, j, i) / S(aj, k, j, i) * c->dt / COEF_TIME_ACCURACY;
  c->rhs_sum = sum;
}

/* ------------------------------------------------------------------ */
/* nullspace                                                            */
/* ------------------------------------------------------------------ */
/* FullyBlocked (poisson.c:1578-1660): KSKE[2*(j*mx+i)] = first solid k of
 * each (i,j) line; multinullspace if every interior line is blocked. */
static void fully_blocked(olevel *L) {
  const int mx = L->mx, my = L->my, mz = L->mz;
  int all = 1;
  for (int j = 1; j < my - 1; j++)
    for (int i = 1; i < mx - 1; i++) {
      int blocked = 0;
      for (int k = 0; k < mz; k++)
        if (S(&L->nvert, k, j, i) > 0.1) {
          if (!blocked) { L->kske[2 * (j * mx + i)] = k; blocked = 1; }
          else if (k < L->kske[2 * (j * mx + i)]) L->kske[2 * (j * mx + i)] = k;
        }
      if (!blocked) all = 0;
    }
  L->multinullspace = all;
}

void o_nullspace(octx *c, int level, double *x) {
  olevel *L = &c->lev[level];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const long N = (long)mx * my * mz;
  const ofield *nv = &L->nvert;
  /* MatNullSpaceRemove with has_cnst: subtract the mean of all entries */
  double sumall = 0.;
  for (long r = 0; r < N; r++) sumall += x[r];
  double shift = -sumall / (double)N;
  for (long r = 0; r < N; r++) x[r] += shift;
  /* PoissonNullSpaceFunction (poisson.c:411-558) */
#define XI(i, j, k) x[(i) + (long)mx * ((j) + (long)my * (k))]
  if (!L->multinullspace) {
    double lsum = 0., lnum = 0.;
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
          if (S(nv, k, j, i) < 0.1) { lsum += XI(i, j, k); lnum++; }
    double sum = lsum / (-1.0 * lnum);
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
          if (S(nv, k, j, i) < 0.1) XI(i, j, k) += sum;
  } else {
    for (int part = 0; part < 2; part++) {
      double lsum = 0., lnum = 0.;
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
          for (int k = 1; k < mz - 1; k++) {
            int below = k < L->kske[2 * (j * mx + i)];
            if ((part == 0 ? below : !below) && S(nv, k, j, i) < 0.1) { lsum += XI(i, j, k); lnum++; }
          }
      double sum = lsum / -lnum;
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
          for (int k = 1; k < mz - 1; k++) {
            int below = k < L->kske[2 * (j * mx + i)];
            if ((part == 0 ? below : !below) && S(nv, k, j, i) < 0.1) XI(i, j, k) += sum;
          }
    }
  }
  for (int j = 0; j < my; j++) for (int i = 0; i < mx; i++) { XI(i, j, 0) = 0.; XI(i, j, mz - 1) = 0.; }
  for (int k = 0; k < mz; k++) for (int i = 0; i < mx; i++) { XI(i, 0, k) = 0.; XI(i, my - 1, k) = 0.; }
  for (int k = 0; k < mz; k++) for (int j = 0; j < my; j++) { XI(0, j, k) = 0.; XI(mx - 1, j, k) = 0.; }
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++)
        if (S(nv, k, j, i) > 0.1) XI(i, j, k) = 0.;
#undef XI
}

/* ------------------------------------------------------------------ */
/* MG transfers                                                         */
/* ------------------------------------------------------------------ */
/* MyRestriction (poisson.c:129-226): lc = coarse level index */
void o_restrict(octx *c, int lc, const double *xf, double *xc) {
  olevel *C = &c->lev[lc], *F = &c->lev[lc + 1];
  const int mx = C->mx, my = C->my, mz = C->mz, fmx = F->mx, fmy = F->my;
  int ia = C->isc ? 0 : 1, ja = C->jsc ? 0 : 1, ka = C->ksc ? 0 : 1;
  const ofield *nf = &F->nvert, *nc = &C->nvert;
#define XF(i, j, k) xf[(i) + (long)fmx * ((j) + (long)fmy * (k))]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        double f;
        if (k == 0 || k == mz - 1 || j == 0 || j == my - 1 || i == 0 || i == mx - 1) f = 0.;
        else {
          int ih = C->isc ? i : 2 * i, jh = C->jsc ? j : 2 * j, kh = C->ksc ? k : 2 * k;
          f = 0.125 * (XF(ih, jh, kh) * fmax(0., 1 - S(nf, kh, jh, ih)) +
                       XF(ih - ia, jh, kh) * fmax(0., 1 - S(nf, kh, jh, ih - ia)) +
                       XF(ih, jh - ja, kh) * fmax(0., 1 - S(nf, kh, jh - ja, ih)) +
                       XF(ih, jh, kh - ka) * fmax(0., 1 - S(nf, kh - ka, jh, ih)) +
                       XF(ih - ia, jh - ja, kh) * fmax(0., 1 - S(nf, kh, jh - ja, ih - ia)) +
                       XF(ih, jh - ja, kh - ka) * fmax(0., 1 - S(nf, kh - ka, jh - ja, ih)) +
                       XF(ih - ia, jh, kh - ka) * fmax(0., 1 - S(nf, kh - ka, jh, ih - ia)) +
                       XF(ih - ia, jh - ja, kh - ka) * fmax(0., 1 - S(nf, kh - ka, jh - ja, ih - ia)));
          if (S(nc, k, j, i) > 0.1) f = 0.;
        }
        xc[i + (long)mx * (j + (long)my * k)] = f;
      }
#undef XF
}

/* MyInterpolation (poisson.c:265-370): lf = fine level index */
void o_interp(octx *c, int lf, const double *xc, double *xf) {
  olevel *F = &c->lev[lf], *C = &c->lev[lf - 1];
  const int mx = F->mx, my = F->my, mz = F->mz, cmx = C->mx, cmy = C->my;
  const ofield *nvc = &C->nvert, *nvf = &F->nvert;
#define XC(i, j, k) xc[(i) + (long)cmx * ((j) + (long)cmy * (k))]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        double f = 0.;
        if (i > 0 && i < mx - 1 && j > 0 && j < my - 1 && k > 0 && k < mz - 1) {
          int ic, jc, kc, ia, ja, ka;
          if (C->isc) { ic = i; ia = 0; }
          else { ic = (i + 1) / 2; ia = (i - 2 * ic) == 0 ? 1 : -1; if (i == 1 || i == mx - 2) ia = 0; }
          if (C->jsc) { jc = j; ja = 0; }
          else { jc = (j + 1) / 2; ja = (j - 2 * jc) == 0 ? 1 : -1; if (j == 1 || j == my - 2) ja = 0; }
          if (C->ksc) { kc = k; ka = 0; }
          else { kc = (k + 1) / 2; ka = (k - 2 * kc) == 0 ? 1 : -1; if (k == 1 || k == mz - 2) ka = 0; }
          if (ka == -1 && S(nvc, kc - 1, jc, ic) > 0.1) ka = 0;
          else if (ka == 1 && S(nvc, kc + 1, jc, ic) > 0.1) ka = 0;
          if (ja == -1 && S(nvc, kc, jc - 1, ic) > 0.1) ja = 0;
          else if (ja == 1 && S(nvc, kc, jc + 1, ic) > 0.1) ja = 0;
          if (ia == -1 && S(nvc, kc, jc, ic - 1) > 0.1) ia = 0;
          else if (ia == 1 && S(nvc, kc, jc, ic + 1) > 0.1) ia = 0;
          f = (XC(ic, jc, kc) * 9 + XC(ic, jc + ja, kc) * 3 + XC(ic + ia, jc, kc) * 3 +
               XC(ic + ia, jc + ja, kc)) * 3. / 64. +
              (XC(ic, jc, kc + ka) * 9 + XC(ic, jc + ja, kc + ka) * 3 + XC(ic + ia, jc, kc + ka) * 3 +
               XC(ic + ia, jc + ja, kc + ka)) / 64.;
        }
        if (S(nvf, k, j, i) > 0.1) f = 0.;
        xf[i + (long)mx * (j + (long)my * k)] = f;
      }
#undef XC
}

/* ------------------------------------------------------------------ */
/* block-Jacobi ILU(0)                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
  int nb;                 /* number of blocks */
  long *r0;               /* block row ranges [r0[b], r0[b+1]) */
  int *rp; int *cl; double *lu; int *diag;  /* per-row restricted pattern */
} objilu;

static void bjilu_build(const olevel *L, int nblocks, objilu *P) {
  const long N = L->nrows;
  if (nblocks < 1) nblocks = 1;
  if (nblocks > N) nblocks = (int)N;
  P->nb = nblocks;
  P->r0 = (long *)malloc((nblocks + 1) * sizeof(long));
  for (int b = 0; b <= nblocks; b++) P->r0[b] = N * b / nblocks;
  P->rp = (int *)malloc((N + 1) * sizeof(int));
  P->rp[0] = 0;
  int *blk = (int *)malloc(N * sizeof(int));
  for (int b = 0; b < nblocks; b++) for (long r = P->r0[b]; r < P->r0[b + 1]; r++) blk[r] = b;
  for (long r = 0; r < N; r++) {
    int n = 0;
    for (int q = L->rowptr[r]; q < L->rowptr[r + 1]; q++)
      if (blk[L->col[q]] == blk[r]) n++;
    P->rp[r + 1] = P->rp[r] + n;
  }
  P->cl = (int *)malloc(P->rp[N] * sizeof(int));
  P->lu = (double *)malloc(P->rp[N] * sizeof(double));
  P->diag = (int *)malloc(N * sizeof(int));
  for (long r = 0; r < N; r++) {
    int n = P->rp[r];
    P->diag[r] = -1;
    for (int q = L->rowptr[r]; q < L->rowptr[r + 1]; q++)
      if (blk[L->col[q]] == blk[r]) {
        if (L->col[q] == r) P->diag[r] = n;
        P->cl[n] = L->col[q]; P->lu[n] = L->val[q]; n++;
      }
  }
  free(blk);
  /* ILU(0) factorisation inside each block (IKJ order); zero pivots are
   * shifted by 1e-10 as PCFactorSetShiftAmount (poisson.c:1896-1910) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nblocks; b++) {
    for (long r = P->r0[b]; r < P->r0[b + 1]; r++) {
      for (int q = P->rp[r]; q < P->diag[r]; q++) {
        int kcol = P->cl[q];
        double piv = P->lu[P->diag[kcol]];
        if (piv == 0.) piv = 1e-10;
        double lik = P->lu[q] / piv;
        P->lu[q] = lik;
        /* a_rj -= lik * a_kj for j>kcol present in both patterns */
        int qa = q + 1;
        for (int qk = P->diag[kcol] + 1; qk < P->rp[kcol + 1]; qk++) {
          int jcol = P->cl[qk];
          while (qa < P->rp[r + 1] && P->cl[qa] < jcol) qa++;
          if (qa < P->rp[r + 1] && P->cl[qa] == jcol) P->lu[qa] -= lik * P->lu[qk];
        }
      }
      if (P->lu[P->diag[r]] == 0.) P->lu[P->diag[r]] = 1e-10;
    }
  }
}

static void bjilu_free(objilu *P) { free(P->r0); free(P->rp); free(P->cl); free(P->lu); free(P->diag); }

/* z = M^{-1} r */
static void bjilu_apply(const objilu *P, const double *r, double *z) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < P->nb; b++) {
    for (long i = P->r0[b]; i < P->r0[b + 1]; i++) {
      double s = r[i];
      for (int q = P->rp[i]; q < P->diag[i]; q++) s -= P->lu[q] * z[P->cl[q]];
      z[i] = s;
    }
    for (long i = P->r0[b + 1] - 1; i >= P->r0[b]; i--) {
      double s = z[i];
      for (int q = P->diag[i] + 1; q < P->rp[i + 1]; q++) s -= P->lu[q] * z[P->cl[q]];
      z[i] = s / P->lu[P->diag[i]];
    }
  }
}

/* ------------------------------------------------------------------ */
/* PCMG V-cycle                                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  objilu pc[O_MAXLEV];
  double *x[O_MAXLEV], *b[O_MAXLEV], *r[O_MAXLEV], *t[O_MAXLEV];
} omg;

static double dot(const double *a, const double *b, long n) {
  double s = 0.;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* Richardson(nit) with left block-Jacobi ILU preconditioner and the
 * level's nullspace removed from each preconditioned correction (the KSP
 * nullspace handling of left preconditioning). */
static void richardson(octx *c, omg *M, int l, int nit) {
  olevel *L = &c->lev[l];
  long N = L->nrows;
  double *x = M->x[l], *b = M->b[l], *r = M->r[l], *t = M->t[l];
  for (int it = 0; it < nit; it++) {
    o_csr_matvec(L, x, r);
    for (long q = 0; q < N; q++) r[q] = b[q] - r[q];
    bjilu_apply(&M->pc[l], r, t);
    o_nullspace(c, l, t);
    for (long q = 0; q < N; q++) x[q] += t[q];
  }
}

static void vcycle(octx *c, omg *M, int l) {
  olevel *L = &c->lev[l];
  long N = L->nrows;
  if (l == 0) {   /* coarse: KSPPREONLY + PCBJACOBI */
    bjilu_apply(&M->pc[0], M->b[0], M->x[0]);
    o_nullspace(c, 0, M->x[0]);
    return;
  }
  memset(M->x[l], 0, N * sizeof(double));
  richardson(c, M, l, c->smooth_it);
  o_csr_matvec(L, M->x[l], M->r[l]);
  for (long q = 0; q < N; q++) M->r[l][q] = M->b[l][q] - M->r[l][q];
  o_restrict(c, l - 1, M->r[l], M->b[l - 1]);
  memset(M->x[l - 1], 0, c->lev[l - 1].nrows * sizeof(double));
  vcycle(c, M, l - 1);
  o_interp(c, l, M->x[l - 1], M->t[l]);
  for (long q = 0; q < N; q++) M->x[l][q] += M->t[l][q];
  richardson(c, M, l, c->smooth_it);
}

/* ------------------------------------------------------------------ */
/* PoissonSolver_MG (poisson.c:1775-2017)                              */
/* ------------------------------------------------------------------ */
int o_poisson_solve(octx *c) {
  const int top = c->nlev - 1;
  olevel *F = &c->lev[top];
  /* nvert hierarchy and multi-nullspace check (poisson.c:1793-1807) */
  o_nvert_restrict_all(c);
  for (int l = top; l > 0; l--) c->lev[l].multinullspace = 0;
  fully_blocked(&c->lev[0]);
  o_volumeflux_rev(c);
  o_poisson_rhs(c, &c->b);
  for (int l = top; l >= 0; l--) o_poisson_lhs(c, l);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  c->nthreads = nth;
  omg M;
  for (int l = 0; l <= top; l++) {
    long N = c->lev[l].nrows;
    bjilu_build(&c->lev[l], nth, &M.pc[l]);
    M.x[l] = (double *)calloc(N, sizeof(double)); M.b[l] = (double *)calloc(N, sizeof(double));
    M.r[l] = (double *)calloc(N, sizeof(double)); M.t[l] = (double *)calloc(N, sizeof(double));
  }
  const long N = F->nrows;
  const int m = c->restart;
  double *b = (double *)malloc(N * sizeof(double));
  o_get_field(&c->b, b);
  double *x = (double *)calloc(N, sizeof(double));
  double *vv = (double *)malloc((size_t)(m + 1) * N * sizeof(double));
  double *zz = (double *)malloc((size_t)m * N * sizeof(double));
  double *w = (double *)malloc(N * sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(m * sizeof(double)), *sn = (double *)malloc(m * sizeof(double));
  double *gv = (double *)malloc((m + 1) * sizeof(double)), *yv = (double *)malloc(m * sizeof(double));
  double bnorm = sqrt(dot(b, b, N));
  c->ksp_bnorm = bnorm;
  double tol = fmax(c->rtol * bnorm, c->atol);
  int its = 0; double rnorm = bnorm;
  if (bnorm == 0.) { rnorm = 0.; }
  while (rnorm > tol && its < c->max_it) {
    /* r = b - A x */
    o_csr_matvec(F, x, w);
    for (long q = 0; q < N; q++) w[q] = b[q] - w[q];
    double beta = sqrt(dot(w, w, N));
    rnorm = beta;
    if (beta <= tol) break;
    for (long q = 0; q < N; q++) vv[q] = w[q] / beta;
    for (int q = 0; q <= m; q++) gv[q] = 0.;
    gv[0] = beta;
    int jj = 0;
    for (; jj < m && its < c->max_it; jj++) {
      /* z_j = M^{-1} v_j */
      memcpy(M.b[top], vv + (size_t)jj * N, N * sizeof(double));
      vcycle(c, &M, top);
      memcpy(zz + (size_t)jj * N, M.x[top], N * sizeof(double));
      o_csr_matvec(F, zz + (size_t)jj * N, w);
      for (int q = 0; q <= jj; q++) {          /* modified Gram-Schmidt */
        double h = dot(w, vv + (size_t)q * N, N);
        H[q * m + jj] = h;
        for (long s = 0; s < N; s++) w[s] -= h * vv[(size_t)q * N + s];
      }
      double hn = sqrt(dot(w, w, N));
      H[(jj + 1) * m + jj] = hn;
      if (hn != 0.) for (long s = 0; s < N; s++) vv[(size_t)(jj + 1) * N + s] = w[s] / hn;
      for (int q = 0; q < jj; q++) {            /* apply previous rotations */
        double a = H[q * m + jj], bb = H[(q + 1) * m + jj];
        H[q * m + jj] = cs[q] * a + sn[q] * bb;
        H[(q + 1) * m + jj] = -sn[q] * a + cs[q] * bb;
      }
      double a = H[jj * m + jj], bb = H[(jj + 1) * m + jj];
      double den = sqrt(a * a + bb * bb);
      cs[jj] = den == 0. ? 1. : a / den; sn[jj] = den == 0. ? 0. : bb / den;
      H[jj * m + jj] = cs[jj] * a + sn[jj] * bb; H[(jj + 1) * m + jj] = 0.;
      gv[jj + 1] = -sn[jj] * gv[jj]; gv[jj] = cs[jj] * gv[jj];
      its++;
      rnorm = fabs(gv[jj + 1]);
      if (rnorm <= tol || hn == 0.) { jj++; break; }
    }
    /* x += Z y */
    for (int q = jj - 1; q >= 0; q--) {
      double s = gv[q];
      for (int t = q + 1; t < jj; t++) s -= H[q * m + t] * yv[t];
      yv[q] = s / H[q * m + q];
    }
    for (int q = 0; q < jj; q++)
      for (long s = 0; s < N; s++) x[s] += yv[q] * zz[(size_t)q * N + s];
    /* true residual for the test */
    o_csr_matvec(F, x, w);
    for (long q = 0; q < N; q++) w[q] = b[q] - w[q];
    rnorm = sqrt(dot(w, w, N));
  }
  c->ksp_its = its; c->ksp_rnorm = rnorm;
  o_nullspace(c, top, x);   /* normalise Phi by the nullspace rule */
  o_put_field(&c->phi, x);
  o_g2l(&c->phi, F->per);   /* lPhi (poisson.c:1985-1986) */
  free(b); free(x); free(vv); free(zz); free(w); free(H); free(cs); free(sn); free(gv); free(yv);
  for (int l = 0; l <= top; l++) {
    bjilu_free(&M.pc[l]); free(M.x[l]); free(M.b[l]); free(M.r[l]); free(M.t[l]);
  }
  return (rnorm <= tol) ? 0 : 1;
}

/* ------------------------------------------------------------------ */
/* UpdatePressure (poisson.c:2908-3248)                                 */
/* ------------------------------------------------------------------ */
void o_update_pressure(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz, IM = L->IM, KM = L->KM, JM = L->JM;
  ofield *p = &c->p, *phi = &c->phi;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) S(p, k, j, i) += S(phi, k, j, i);
  if (!(L->bctype[0] == 7 || L->bctype[4] == 7)) return;
  o_g2l(p, L->per); o_g2l(phi, L->per);
  /* first pass, interior of the other axes */
  if (L->bctype[0] == 7) {
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        if (k > 0 && k < KM && j > 0 && j < JM) {
          S(p,k,j,0) = S(p,k,j,-2); S(phi,k,j,0) = S(phi,k,j,-2);
          S(p,k,j,mx-1) = S(p,k,j,mx+1); S(phi,k,j,mx-1) = S(phi,k,j,mx+1);
        }
  }
  o_g2l(p, L->per); o_g2l(phi, L->per);
  /* (j periodic branch not restated) */
  o_g2l(p, L->per); o_g2l(phi, L->per);
  if (L->bctype[4] == 7) {
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++)
        if (i > 0 && i < IM && j > 0 && j < JM) {
          S(p,0,j,i) = S(p,-2,j,i); S(phi,0,j,i) = S(phi,-2,j,i);
          S(p,mz-1,j,i) = S(p,mz+1,j,i); S(phi,mz-1,j,i) = S(phi,mz+1,j,i);
        }
  }
  o_g2l(p, L->per); o_g2l(phi, L->per);
  /* second pass, full planes */
  if (L->bctype[0] == 7) {
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++) { S(p,k,j,0) = S(p,k,j,-2); S(phi,k,j,0) = S(phi,k,j,-2); }
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++) { S(p,k,j,mx-1) = S(p,k,j,mx+1); S(phi,k,j,mx-1) = S(phi,k,j,mx+1); }
  }
  o_g2l(p, L->per); o_g2l(phi, L->per);
  o_g2l(p, L->per); o_g2l(phi, L->per);
  if (L->bctype[4] == 7) {
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) { S(p,mz-1,j,i) = S(p,mz+1,j,i); S(phi,mz-1,j,i) = S(phi,mz+1,j,i); }
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) { S(p,0,j,i) = S(p,-2,j,i); S(phi,0,j,i) = S(phi,-2,j,i); }
  }
  o_g2l(p, L->per); o_g2l(phi, L->per);
}

/* ------------------------------------------------------------------ */
/* Projection (poisson.c:2281-2906)                                     */
/* ------------------------------------------------------------------ */
/* tangential derivative of a scalar field on a face (the A.3 rule with
 * DROP -> 0).  Face between cell (i,j,k) and (i,j,k)+o (o along d). */
double o_face_tangent(const olevel *L, const ofield *p, int d, int t, int i, int j, int k, int *mode_out) {
  const ofield *nv = &L->nvert;
  int o[3] = {0, 0, 0}; o[d] = 1;
  int up[3] = {0, 0, 0}, dn[3] = {0, 0, 0}; up[t] = 1; dn[t] = -1;
  int m_ax[3] = {L->mx, L->my, L->mz}, ijk[3] = {i, j, k};
  double hi = S(nv, k + up[2], j + up[1], i + up[0]) + S(nv, k + up[2] + o[2], j + up[1] + o[1], i + up[0] + o[0]);
  double lo = S(nv, k + dn[2], j + dn[1], i + dn[0]) + S(nv, k + dn[2] + o[2], j + dn[1] + o[1], i + dn[0] + o[0]);
  int per = (t == 0) ? (L->bctype[0] == 7) : (t == 1) ? (L->bctype[2] == 7) : (L->bctype[4] == 7);
  int mode = o_xmode(hi, lo, ijk[t] == m_ax[t] - 2, ijk[t] == 1, per);
  if (mode_out) *mode_out = mode;
#define PP(di, dj, dk) S(p, k + (dk), j + (dj), i + (di))
  if (mode == XM_LOWER)
    return (PP(0,0,0) + PP(o[0],o[1],o[2]) - PP(dn[0],dn[1],dn[2]) - PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.5;
  if (mode == XM_UPPER)
    return (PP(up[0],up[1],up[2]) + PP(o[0]+up[0],o[1]+up[1],o[2]+up[2]) - PP(0,0,0) - PP(o[0],o[1],o[2])) * 0.5;
  if (mode == XM_CENTRAL)
    return (PP(up[0],up[1],up[2]) + PP(o[0]+up[0],o[1]+up[1],o[2]+up[2]) - PP(dn[0],dn[1],dn[2]) - PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.25;
#undef PP
  return 0.;
}

static void project_face(octx *c, int d, int i, int j, int k) {
  olevel *L = &c->lev[c->nlev - 1];
  const ofield *p = &c->phi, *nv = &L->nvert;
  double dp[3];
  int o[3] = {0, 0, 0}; o[d] = 1;
  for (int t = 0; t < 3; t++) {
    if (t == d) dp[t] = S(p, k + o[2], j + o[1], i + o[0]) - S(p, k, j, i);
    else dp[t] = o_face_tangent(L, p, d, t, i, j, k, NULL);
  }
  if (!(S(nv, k, j, i) + S(nv, k + o[2], j + o[1], i + o[0]))) {
    double g[3];
    gcoef(L, d, i, j, k, g);
    double corr = (dp[0] * g[0] + dp[1] * g[1] + dp[2] * g[2]) * c->dt * c->st / COEF_TIME_ACCURACY;
    ocmp *u = &V(&c->ucont, k, j, i);
    if (d == 0) u->x -= corr; else if (d == 1) u->y -= corr; else u->z -= corr;
  }
}

void o_projection(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  int i_end = L->bctype[0] == 7 ? mx - 1 : mx - 2;
  int j_end = my - 2;
  int k_end = L->bctype[4] == 7 ? mz - 1 : mz - 2;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        if (i < i_end) project_face(c, 0, i, j, k);
        if (j < j_end) project_face(c, 1, i, j, k);
        if (k < k_end) project_face(c, 2, i, j, k);
      }
  if (L->bctype[0] == 7)
    for (int k = 1; k < mz - 1; k++) for (int j = 1; j < my - 1; j++) project_face(c, 0, 0, j, k);
  if (L->bctype[4] == 7)
    for (int j = 1; j < my - 1; j++) for (int i = 1; i < mx - 1; i++) project_face(c, 2, i, j, 0);
  o_g2l(&c->ucont, L->per);
  o_contra2cart(c);
  o_ghostnodevelocity(c);
}

/* ------------------------------------------------------------------ */
/* VolumeFlux_rev (poisson.c:3712-3944)                                 */
/* ------------------------------------------------------------------ */
static double cnorm(const ocmp *a) { return sqrt(a->x * a->x + a->y * a->y + a->z * a->z); }

static void ibm_flux_sum(octx *c, double *flux, double *area) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const ofield *nv = &L->nvert; const double ibmval = 1.5;
  double f = 0., a = 0.;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        const ocmp *u = &V(&c->ucont, k, j, i);
        double n0 = S(nv, k, j, i);
        if (n0 < 0.1) {
          if (S(nv,k,j,i+1) > ibmval - 0.4 && S(nv,k,j,i+1) < ibmval && i < mx - 2) { f += u->x; a += cnorm(&V(&L->csi,k,j,i)); }
          if (S(nv,k,j+1,i) > ibmval - 0.4 && S(nv,k,j+1,i) < ibmval && j < my - 2) { f += u->y; a += cnorm(&V(&L->eta,k,j,i)); }
          if (S(nv,k+1,j,i) > ibmval - 0.4 && S(nv,k+1,j,i) < ibmval && k < mz - 2) { f += u->z; a += cnorm(&V(&L->zet,k,j,i)); }
        }
        if (n0 > ibmval - 0.4 && n0 < ibmval) {
          if (S(nv,k,j,i+1) < 0.1 && i < mx - 2) { f -= u->x; a += cnorm(&V(&L->csi,k,j,i)); }
          if (S(nv,k,j+1,i) < 0.1 && j < my - 2) { f -= u->y; a += cnorm(&V(&L->eta,k,j,i)); }
          if (S(nv,k+1,j,i) < 0.1 && k < mz - 2) { f -= u->z; a += cnorm(&V(&L->zet,k,j,i)); }
        }
      }
  *flux = f; *area = a;
}

void o_volumeflux_rev(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const ofield *nv = &L->nvert; const double ibmval = 1.5;
  double flux, area;
  ibm_flux_sum(c, &flux, &area);
  c->ibm_flux0 = flux; c->ibm_area0 = area;
  double corr = area > 1.e-15 ? (flux + 0.0 /* FluxIntpSum */) / area : 0.;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ocmp *u = &V(&c->ucont, k, j, i);
        double n0 = S(nv, k, j, i);
        if (n0 < 0.1) {
          if (S(nv,k,j,i+1) > ibmval - 0.4 && S(nv,k,j,i+1) < ibmval && i < mx - 2) u->x -= cnorm(&V(&L->csi,k,j,i)) * corr;
          if (S(nv,k,j+1,i) > ibmval - 0.4 && S(nv,k,j+1,i) < ibmval && j < my - 2) u->y -= cnorm(&V(&L->eta,k,j,i)) * corr;
          if (S(nv,k+1,j,i) > ibmval - 0.4 && S(nv,k+1,j,i) < ibmval && k < mz - 2) u->z -= cnorm(&V(&L->zet,k,j,i)) * corr;
        }
        if (n0 > ibmval - 0.4 && n0 < ibmval) {
          if (S(nv,k,j,i+1) < 0.1 && i < mx - 2) u->x += cnorm(&V(&L->csi,k,j,i)) * corr;
          if (S(nv,k,j+1,i) < 0.1 && j < my - 2) u->y += cnorm(&V(&L->eta,k,j,i)) * corr;
          if (S(nv,k+1,j,i) < 0.1 && k < mz - 2) u->z += cnorm(&V(&L->zet,k,j,i)) * corr;
        }
      }
  ibm_flux_sum(c, &flux, &area);
  c->ibm_flux1 = flux; c->ibm_area1 = area;
  o_g2l(&c->ucont, L->per);
}

/* ------------------------------------------------------------------ */
/* Divergence (main.c:469-593)                                          */
/* ------------------------------------------------------------------ */
double o_divergence(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const ofield *nv = &L->nvert, *uc = &c->ucont, *aj = &L->aj;
  double best = -1.0; long bidx = 0;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        double dv = 0.;
        if (i > 0 && i < mx - 1 && j > 0 && j < my - 1 && k > 0 && k < mz - 1) {
          dv = fabs((V(uc,k,j,i).x - V(uc,k,j,i-1).x + V(uc,k,j,i).y - V(uc,k,j-1,i).y +
                     V(uc,k,j,i).z - V(uc,k-1,j,i).z) * S(aj, k, j, i));
          if (S(nv,k,j,i) + S(nv,k+1,j,i) + S(nv,k-1,j,i) + S(nv,k,j+1,i) + S(nv,k,j-1,i) +
              S(nv,k,j,i+1) + S(nv,k,j,i-1) > 0.1) dv = 0.;
        }
        if (dv > best) { best = dv; bidx = i + (long)mx * (j + (long)my * k); }
      }
  c->maxdiv = best; c->maxdiv_idx = bidx;
  return best;
}
