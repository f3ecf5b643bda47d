/*
 * cv_oracle_grid.c -- TEST INFRASTRUCTURE (CPU oracle): padded DMDA-like
 * fields, grid hierarchy, FormMetrics + Update_Metrics_PBC, nvert
 * restriction.
 *
 * Parity pins:
 *   FormMetrics        closed forms on uniform and affine grids (face vectors
 *                      = lattice cofactors, 1/J = triple product;
 *                      tests/test_oracle_pins.py); the geometric conservation
 *                      law of MetricsDivergence (metrics.c:1551-1625), face
 *                      vectors of every cell summing to 0 within 1e-13 on the
 *                      warped channel and on jittered nodes
 *                      (tests/test_oracle_grid.py).
 *   MyNvertRestriction a direct numpy evaluation of the two passes on random
 *                      masks, bit for bit on every level, full and
 *                      semi-coarsened (tests/test_oracle_grid.py).
 * Only single-rank semantics are restated (xs=0, xe=mx, ...).
 * Supported boundary types: 1 (no-slip wall) on every axis, 7 (periodic)
 * on the i and k axes, 3 (slip) and 6 (far field) on every face but j-, 5 (inlet)
 * on the k- face and 4 (outlet) on the k+ face; metrics and the Poisson
 * operator only tell 7 from the others.  The cgrid option is taken as 0 (the reference
 * leaves it uninitialised unless -cgrid is given, init.c:356-358).
 */
#include "cv_oracle_int.h"
#include <stdio.h>

/* ------------------------------------------------------------------ */
/* fields                                                               */
/* ------------------------------------------------------------------ */
void of_alloc(ofield *f, int mx, int my, int mz, int dof) {
  f->m[0] = mx; f->m[1] = my; f->m[2] = mz;
  f->p[0] = mx + 2 * OG; f->p[1] = my + 2 * OG; f->p[2] = mz + 2 * OG;
  f->dof = dof;
  f->a = (double *)calloc((size_t)f->p[0] * f->p[1] * f->p[2] * dof, sizeof(double));
  if (!f->a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
}
void of_free(ofield *f) { free(f->a); f->a = NULL; }
void of_zero(ofield *f) {
  memset(f->a, 0, (size_t)f->p[0] * f->p[1] * f->p[2] * f->dof * sizeof(double));
}
void of_copy(ofield *dst, const ofield *src) {
  memcpy(dst->a, src->a, (size_t)src->p[0] * src->p[1] * src->p[2] * src->dof * sizeof(double));
}

/* DMGlobalToLocal for a single rank: ghost points of periodic axes take the
 * value of global point (l mod m); ghost points outside a non-periodic axis
 * do not exist in PETSc and are left untouched here. */
void o_g2l(ofield *f, const int per[3]) {
  const int mx = f->m[0], my = f->m[1], mz = f->m[2], d = f->dof;
  for (int k = -OG; k < mz + OG; k++) {
    int kin = (k >= 0 && k < mz);
    if (!kin && !per[2]) continue;
    int kk = ((k % mz) + mz) % mz;
    for (int j = -OG; j < my + OG; j++) {
      int jin = (j >= 0 && j < my);
      if (!jin && !per[1]) continue;
      int jj = ((j % my) + my) % my;
      for (int i = -OG; i < mx + OG; i++) {
        int iin = (i >= 0 && i < mx);
        if (!iin && !per[0]) continue;
        if (iin && jin && kin) continue;
        int ii = ((i % mx) + mx) % mx;
        size_t dst = oix(f, i, j, k) * d, src = oix(f, ii, jj, kk) * d;
        for (int c = 0; c < d; c++) f->a[dst + c] = f->a[src + c];
      }
    }
  }
}

void o_get_field(const ofield *f, double *out) {
  const int d = f->dof; size_t n = 0;
  for (int k = 0; k < f->m[2]; k++)
    for (int j = 0; j < f->m[1]; j++)
      for (int i = 0; i < f->m[0]; i++) {
        size_t s = oix(f, i, j, k) * d;
        for (int c = 0; c < d; c++) out[n++] = f->a[s + c];
      }
}
void o_put_field(ofield *f, const double *in) {
  const int d = f->dof; size_t n = 0;
  for (int k = 0; k < f->m[2]; k++)
    for (int j = 0; j < f->m[1]; j++)
      for (int i = 0; i < f->m[0]; i++) {
        size_t s = oix(f, i, j, k) * d;
        for (int c = 0; c < d; c++) f->a[s + c] = in[n++];
      }
}

/* ------------------------------------------------------------------ */
/* hierarchy: MGDACreate dimensions (init.c:70-107)                     */
/* ------------------------------------------------------------------ */
static void level_alloc(olevel *L) {
  int mx = L->mx, my = L->my, mz = L->mz;
  ofield *v3[] = {&L->coor, &L->csi, &L->eta, &L->zet, &L->icsi, &L->ieta, &L->izet,
                  &L->jcsi, &L->jeta, &L->jzet, &L->kcsi, &L->keta, &L->kzet,
                  &L->cent, &L->centx, &L->centy, &L->centz, &L->gs};
  for (size_t q = 0; q < sizeof(v3) / sizeof(v3[0]); q++) of_alloc(v3[q], mx, my, mz, 3);
  ofield *v1[] = {&L->aj, &L->iaj, &L->jaj, &L->kaj, &L->nvert};
  for (size_t q = 0; q < sizeof(v1) / sizeof(v1[0]); q++) of_alloc(v1[q], mx, my, mz, 1);
  L->rowptr = NULL; L->col = NULL; L->val = NULL; L->nrows = 0;
  L->kske = (int *)calloc((size_t)2 * mx * my, sizeof(int));
  L->multinullspace = 0;
}

static void level_free(olevel *L) {
  ofield *v[] = {&L->coor, &L->csi, &L->eta, &L->zet, &L->icsi, &L->ieta, &L->izet,
                 &L->jcsi, &L->jeta, &L->jzet, &L->kcsi, &L->keta, &L->kzet,
                 &L->cent, &L->centx, &L->centy, &L->centz, &L->gs,
                 &L->aj, &L->iaj, &L->jaj, &L->kaj, &L->nvert};
  for (size_t q = 0; q < sizeof(v) / sizeof(v[0]); q++) of_free(v[q]);
  free(L->rowptr); free(L->col); free(L->val); free(L->kske);
}

/* coarsen_auto: at each level transition coarsen, among the axes the semi
 * flags allow, only those with the most cells (per-level flags, stored on
 * the coarser level of each transition); 0: MGDACreate's fixed flags */
octx *o_create(int IM, int JM, int KM, const int bctype[6], int nlev,
               int isc, int jsc, int ksc, double ren, double dt, int coarsen_auto) {
  if (nlev < 1 || nlev > O_MAXLEV) return NULL;
  if (bctype[2] == 7 || bctype[3] == 7) return NULL;  /* j-periodic not restated */
  /* every face: wall 1, slip 3, far field 6, periodic 7 (not on j); k-
   * also inlet 5, k+ also outlet 4.  Not slip or far field on j-:
   * FormMetrics leaves eta of the j = 0 face at 0 (its loop starts at
   * j = 1, metrics.c:107-110), so both divide by zero there. */
  for (int q = 0; q < 6; q++) {
    const int b = bctype[q];
    if (q == 2 && (b == 3 || b == 6)) return NULL;
    if (b == 1 || b == 3 || b == 6 || b == 7) continue;
    if ((q == 4 && b == 5) || (q == 5 && b == 4)) continue;
    return NULL;
  }
  if ((bctype[0] == 7) != (bctype[1] == 7) || (bctype[4] == 7) != (bctype[5] == 7)) return NULL;
  octx *c = (octx *)calloc(1, sizeof(octx));
  c->nlev = nlev;
  c->ren = ren; c->dt = dt; c->st = 1.0;          /* init.c:52-62 */
  c->restart = 30; c->max_it = 500; c->rtol = 1e-10; c->atol = 0.0;
  c->smooth_it = 2; c->nthreads = 1; c->formbcs_passes = 3;
  int im = IM, jm = JM, km = KM;
  for (int l = nlev - 1; l >= 0; l--) {
    olevel *L = &c->lev[l];
    int si = isc, sj = jsc, sk = ksc;
    if (coarsen_auto && l < nlev - 1) {
      int most = 0;
      if (!si && im - 1 > most) most = im - 1;
      if (!sj && jm - 1 > most) most = jm - 1;
      if (!sk && km - 1 > most) most = km - 1;
      if (!si && im - 1 < most) si = 1;
      if (!sj && jm - 1 < most) sj = 1;
      if (!sk && km - 1 < most) sk = 1;
    }
    L->isc = si; L->jsc = sj; L->ksc = sk;
    if (l < nlev - 1) {          /* MGDACreate, init.c:78-99 */
      im = si ? im : (im + 1) / 2;
      jm = sj ? jm : (jm + 1) / 2;
      km = sk ? km : (km + 1) / 2;
    }
    L->IM = im; L->JM = jm; L->KM = km;
    L->mx = im + 1; L->my = jm + 1; L->mz = km + 1;
    memcpy(L->bctype, bctype, 6 * sizeof(int));
    L->per[0] = bctype[0] == 7; L->per[1] = bctype[2] == 7; L->per[2] = bctype[4] == 7;
    L->thislevel = l;
    level_alloc(L);
  }
  olevel *F = &c->lev[nlev - 1];
  ofield *v3[] = {&c->ucont, &c->ucont_o, &c->ucat, &c->rhs, &c->ubcs, &c->ucont_rm1, &c->pucont};
  for (size_t q = 0; q < 7; q++) of_alloc(v3[q], F->mx, F->my, F->mz, 3);
  c->mom_scheme = 0; c->imp_max_it = 50; c->imp_atol = 1e-7; c->imp_rtol = 1e-4; c->cfl = 0.2;
  c->inletprofile = 2;                            /* -inlet default (main.c:23) */
  of_alloc(&c->p, F->mx, F->my, F->mz, 1);
  of_alloc(&c->phi, F->mx, F->my, F->mz, 1);
  of_alloc(&c->nvert_o, F->mx, F->my, F->mz, 1);
  of_alloc(&c->b, F->mx, F->my, F->mz, 1);
  return c;
}

void o_destroy(octx *c) {
  if (!c) return;
  for (int l = 0; l < c->nlev; l++) level_free(&c->lev[l]);
  ofield *v[] = {&c->ucont, &c->ucont_o, &c->ucat, &c->rhs, &c->ubcs, &c->p, &c->phi,
                 &c->nvert_o, &c->b, &c->ucont_rm1, &c->pucont};
  for (size_t q = 0; q < sizeof(v) / sizeof(v[0]); q++) of_free(v[q]);
  free(c->ib_idx); free(c->ib_val);
  free(c);
}

/* Node coordinates: grid.dat 3-D order (init.c:689-720): x(k,j,i) for all
 * nodes, then y, then z; here passed interleaved per node.  Storage index
 * IM (=mx-1) keeps coordinate 0 (VecSet 0, init.c:666).  Coarser levels
 * take every second fine node (GridRestriction, init.c:786-828). */
void o_set_coords(octx *c, const double *xyz) {
  olevel *F = &c->lev[c->nlev - 1];
  of_zero(&F->coor);
  for (int k = 0; k < F->KM; k++)
    for (int j = 0; j < F->JM; j++)
      for (int i = 0; i < F->IM; i++) {
        size_t n = ((size_t)k * F->JM + j) * F->IM + i;
        ocmp *p = &V(&F->coor, k, j, i);
        p->x = xyz[3 * n]; p->y = xyz[3 * n + 1]; p->z = xyz[3 * n + 2];
      }
  o_g2l(&F->coor, F->per);
  for (int l = c->nlev - 2; l >= 0; l--) {
    olevel *L = &c->lev[l], *H = &c->lev[l + 1];
    of_zero(&L->coor);
    for (int k = 0; k < L->mz - 1; k++)
      for (int j = 0; j < L->my - 1; j++)
        for (int i = 0; i < L->mx - 1; i++) {
          int ih = L->isc ? i : 2 * i, jh = L->jsc ? j : 2 * j, kh = L->ksc ? k : 2 * k;
          V(&L->coor, k, j, i) = V(&H->coor, kh, jh, ih);
        }
    o_g2l(&L->coor, L->per);
  }
}

/* ------------------------------------------------------------------ */
/* FormMetrics (metrics.c:4-1229), single rank, cgrid = 0               */
/* ------------------------------------------------------------------ */
#define CX(f, k, j, i) V(f, k, j, i).x
#define CY(f, k, j, i) V(f, k, j, i).y
#define CZ(f, k, j, i) V(f, k, j, i).z

static void cross_metric(double dxa, double dya, double dza, double dxb, double dyb, double dzb,
                         ocmp *out) {
  /* out = a x b, written exactly as metrics.c:100-102 */
  out->x = dya * dzb - dza * dyb;
  out->y = -dxa * dzb + dza * dxb;
  out->z = dxa * dyb - dya * dxb;
}

static void form_metrics_level(olevel *L) {
  const int mx = L->mx, my = L->my, mz = L->mz;
  const int IM = L->IM, JM = L->JM, KM = L->KM;
  const int sw = (L->per[0] || L->per[1] || L->per[2]) ? 3 : 2;   /* init.c:110-157 */
  const int gxs = L->per[0] ? -sw : 0, gxe = L->per[0] ? mx + sw : mx;
  const int gys = L->per[1] ? -sw : 0, gye = L->per[1] ? my + sw : my;
  const int gzs = L->per[2] ? -sw : 0, gze = L->per[2] ? mz + sw : mz;
  const int lxs = 1, lxe = mx - 1, lys = 1, lye = my - 1, lzs = 1, lze = mz - 1;
  ofield *co = &L->coor;
  ofield *csi = &L->csi, *eta = &L->eta, *zet = &L->zet, *aj = &L->aj;
  of_zero(csi); of_zero(eta); of_zero(zet); of_zero(aj);
  of_zero(&L->centx); of_zero(&L->centy); of_zero(&L->centz);
  of_zero(&L->cent); of_zero(&L->gs);
  of_zero(&L->icsi); of_zero(&L->ieta); of_zero(&L->izet); of_zero(&L->iaj);
  of_zero(&L->jcsi); of_zero(&L->jeta); of_zero(&L->jzet); of_zero(&L->jaj);
  of_zero(&L->kcsi); of_zero(&L->keta); of_zero(&L->kzet); of_zero(&L->kaj);
  double dxdc, dydc, dzdc, dxde, dyde, dzde, dxdz, dydz, dzdz;

  /* csi = de x dz  (metrics.c:80-106) */
  for (int k = lzs; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = 0; i < lxe; i++) {
        dxde = 0.5 * (CX(co,k,j,i) + CX(co,k-1,j,i) - CX(co,k,j-1,i) - CX(co,k-1,j-1,i));
        dyde = 0.5 * (CY(co,k,j,i) + CY(co,k-1,j,i) - CY(co,k,j-1,i) - CY(co,k-1,j-1,i));
        dzde = 0.5 * (CZ(co,k,j,i) + CZ(co,k-1,j,i) - CZ(co,k,j-1,i) - CZ(co,k-1,j-1,i));
        dxdz = 0.5 * (CX(co,k,j-1,i) + CX(co,k,j,i) - CX(co,k-1,j-1,i) - CX(co,k-1,j,i));
        dydz = 0.5 * (CY(co,k,j-1,i) + CY(co,k,j,i) - CY(co,k-1,j-1,i) - CY(co,k-1,j,i));
        dzdz = 0.5 * (CZ(co,k,j-1,i) + CZ(co,k,j,i) - CZ(co,k-1,j-1,i) - CZ(co,k-1,j,i));
        cross_metric(dxde, dyde, dzde, dxdz, dydz, dzdz, &V(csi, k, j, i));
      }
  /* eta = dz x dc  (metrics.c:109-132) */
  for (int k = lzs; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = lxs; i < lxe; i++) {
        dxdc = 0.5 * (CX(co,k,j,i) + CX(co,k-1,j,i) - CX(co,k,j,i-1) - CX(co,k-1,j,i-1));
        dydc = 0.5 * (CY(co,k,j,i) + CY(co,k-1,j,i) - CY(co,k,j,i-1) - CY(co,k-1,j,i-1));
        dzdc = 0.5 * (CZ(co,k,j,i) + CZ(co,k-1,j,i) - CZ(co,k,j,i-1) - CZ(co,k-1,j,i-1));
        dxdz = 0.5 * (CX(co,k,j,i) + CX(co,k,j,i-1) - CX(co,k-1,j,i) - CX(co,k-1,j,i-1));
        dydz = 0.5 * (CY(co,k,j,i) + CY(co,k,j,i-1) - CY(co,k-1,j,i) - CY(co,k-1,j,i-1));
        dzdz = 0.5 * (CZ(co,k,j,i) + CZ(co,k,j,i-1) - CZ(co,k-1,j,i) - CZ(co,k-1,j,i-1));
        cross_metric(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &V(eta, k, j, i));
      }
  /* zet = dc x de  (metrics.c:135-157); k starts at zs=0 */
  for (int k = 0; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = lxs; i < lxe; i++) {
        dxdc = 0.5 * (CX(co,k,j,i) + CX(co,k,j-1,i) - CX(co,k,j,i-1) - CX(co,k,j-1,i-1));
        dydc = 0.5 * (CY(co,k,j,i) + CY(co,k,j-1,i) - CY(co,k,j,i-1) - CY(co,k,j-1,i-1));
        dzdc = 0.5 * (CZ(co,k,j,i) + CZ(co,k,j-1,i) - CZ(co,k,j,i-1) - CZ(co,k,j-1,i-1));
        dxde = 0.5 * (CX(co,k,j,i) + CX(co,k,j,i-1) - CX(co,k,j-1,i) - CX(co,k,j-1,i-1));
        dyde = 0.5 * (CY(co,k,j,i) + CY(co,k,j,i-1) - CY(co,k,j-1,i) - CY(co,k,j-1,i-1));
        dzde = 0.5 * (CZ(co,k,j,i) + CZ(co,k,j,i-1) - CZ(co,k,j-1,i) - CZ(co,k,j-1,i-1));
        cross_metric(dxdc, dydc, dzdc, dxde, dyde, dzde, &V(zet, k, j, i));
      }
  /* Jacobian (metrics.c:160-212) */
  for (int k = lzs; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = lxs; i < lxe; i++) {
        const ocmp *a = &V(co,k,j,i), *b = &V(co,k,j-1,i), *c2 = &V(co,k-1,j,i), *d = &V(co,k-1,j-1,i);
        const ocmp *e = &V(co,k,j,i-1), *f = &V(co,k,j-1,i-1), *g = &V(co,k-1,j,i-1), *h = &V(co,k-1,j-1,i-1);
        dxdc = 0.25 * (a->x + b->x + c2->x + d->x - e->x - f->x - g->x - h->x);
        dydc = 0.25 * (a->y + b->y + c2->y + d->y - e->y - f->y - g->y - h->y);
        dzdc = 0.25 * (a->z + b->z + c2->z + d->z - e->z - f->z - g->z - h->z);
        dxde = 0.25 * (a->x + e->x + c2->x + g->x - b->x - f->x - d->x - h->x);
        dyde = 0.25 * (a->y + e->y + c2->y + g->y - b->y - f->y - d->y - h->y);
        dzde = 0.25 * (a->z + e->z + c2->z + g->z - b->z - f->z - d->z - h->z);
        dxdz = 0.25 * (a->x + b->x + e->x + f->x - c2->x - d->x - g->x - h->x);
        dydz = 0.25 * (a->y + b->y + e->y + f->y - c2->y - d->y - g->y - h->y);
        dzdz = 0.25 * (a->z + b->z + e->z + f->z - c2->z - d->z - g->z - h->z);
        double det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                     dzdc * (dxde * dydz - dyde * dxdz);
        S(aj, k, j, i) = 1. / det;
      }
  /* boundary copies (metrics.c:215-322) */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++) {
      V(eta,k,j,0) = V(eta,k,j,1); V(zet,k,j,0) = V(zet,k,j,1); S(aj,k,j,0) = S(aj,k,j,1);
      V(eta,k,j,mx-1) = V(eta,k,j,mx-2); V(zet,k,j,mx-1) = V(zet,k,j,mx-2);
      S(aj,k,j,mx-1) = S(aj,k,j,mx-2);
    }
  for (int k = 0; k < mz; k++)
    for (int i = 0; i < mx; i++) {
      V(csi,k,0,i) = V(csi,k,1,i); V(zet,k,0,i) = V(zet,k,1,i); S(aj,k,0,i) = S(aj,k,1,i);
      V(csi,k,my-1,i) = V(csi,k,my-2,i); V(zet,k,my-1,i) = V(zet,k,my-2,i);
      S(aj,k,my-1,i) = S(aj,k,my-2,i);
    }
  for (int j = 0; j < my; j++)
    for (int i = 0; i < mx; i++) {
      V(eta,0,j,i) = V(eta,1,j,i); V(csi,0,j,i) = V(csi,1,j,i); S(aj,0,j,i) = S(aj,1,j,i);
      V(eta,mz-1,j,i) = V(eta,mz-2,j,i); V(csi,mz-1,j,i) = V(csi,mz-2,j,i);
      S(aj,mz-1,j,i) = S(aj,mz-2,j,i);
    }

  /* cell centres and grid spacing (metrics.c:332-425) */
  ofield *cent = &L->cent, *gs = &L->gs;
  for (int k = lzs; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = lxs; i < lxe; i++) {
        ocmp *c8[8] = {&V(co,k,j,i), &V(co,k,j-1,i), &V(co,k-1,j,i), &V(co,k-1,j-1,i),
                       &V(co,k,j,i-1), &V(co,k,j-1,i-1), &V(co,k-1,j,i-1), &V(co,k-1,j-1,i-1)};
        V(cent,k,j,i).x = 0.125 * (c8[0]->x + c8[1]->x + c8[2]->x + c8[3]->x + c8[4]->x + c8[5]->x + c8[6]->x + c8[7]->x);
        V(cent,k,j,i).y = 0.125 * (c8[0]->y + c8[1]->y + c8[2]->y + c8[3]->y + c8[4]->y + c8[5]->y + c8[6]->y + c8[7]->y);
        V(cent,k,j,i).z = 0.125 * (c8[0]->z + c8[1]->z + c8[2]->z + c8[3]->z + c8[4]->z + c8[5]->z + c8[6]->z + c8[7]->z);
        double xcp, ycp, zcp, xcm, ycm, zcm;
        const ocmp *a = c8[0], *b = c8[1], *c2 = c8[2], *d = c8[3];
        const ocmp *e = c8[4], *f = c8[5], *g = c8[6], *h = c8[7];
        /* a=(k,j,i) b=(k,j-1,i) c2=(k-1,j,i) d=(k-1,j-1,i) e..h the same at i-1 */
        xcp = 0.25 * (a->x + b->x + d->x + c2->x); ycp = 0.25 * (a->y + b->y + d->y + c2->y);
        zcp = 0.25 * (a->z + b->z + d->z + c2->z);
        xcm = 0.25 * (e->x + f->x + h->x + g->x); ycm = 0.25 * (e->y + f->y + h->y + g->y);
        zcm = 0.25 * (e->z + f->z + h->z + g->z);
        V(gs,k,j,i).x = sqrt((xcp-xcm)*(xcp-xcm) + (ycp-ycm)*(ycp-ycm) + (zcp-zcm)*(zcp-zcm));
        xcp = 0.25 * (a->x + e->x + c2->x + g->x); ycp = 0.25 * (a->y + e->y + c2->y + g->y);
        zcp = 0.25 * (a->z + e->z + c2->z + g->z);
        xcm = 0.25 * (b->x + f->x + d->x + h->x); ycm = 0.25 * (b->y + f->y + d->y + h->y);
        zcm = 0.25 * (b->z + f->z + d->z + h->z);
        V(gs,k,j,i).y = sqrt((xcp-xcm)*(xcp-xcm) + (ycp-ycm)*(ycp-ycm) + (zcp-zcm)*(zcp-zcm));
        xcp = 0.25 * (a->x + e->x + b->x + f->x); ycp = 0.25 * (a->y + e->y + b->y + f->y);
        zcp = 0.25 * (a->z + e->z + b->z + f->z);
        xcm = 0.25 * (c2->x + g->x + d->x + h->x); ycm = 0.25 * (c2->y + g->y + d->y + h->y);
        zcm = 0.25 * (c2->z + g->z + d->z + h->z);
        V(gs,k,j,i).z = sqrt((xcp-xcm)*(xcp-xcm) + (ycp-ycm)*(ycp-ycm) + (zcp-zcm)*(zcp-zcm));
      }
  o_g2l(gs, L->per);  /* lGridSpace (metrics.c:418-419) */

  /* i-face centres (metrics.c:644-662): i in [0,IM], j in [1,JM], k in [1,KM] */
  ofield *cx = &L->centx;
  for (int k = gzs + 1; k < gze; k++)
    for (int j = gys + 1; j < gye; j++)
      for (int i = gxs; i < gxe; i++) {
        if (!(i > -1 && j > 0 && k > 0 && i < IM + 1 && j < JM + 1 && k < KM + 1)) continue;
        V(cx,k,j,i).x = (CX(co,k,j,i) + CX(co,k-1,j,i) + CX(co,k,j-1,i) + CX(co,k-1,j-1,i)) * 0.25;
        V(cx,k,j,i).y = (CY(co,k,j,i) + CY(co,k-1,j,i) + CY(co,k,j-1,i) + CY(co,k-1,j-1,i)) * 0.25;
        V(cx,k,j,i).z = (CZ(co,k,j,i) + CZ(co,k-1,j,i) + CZ(co,k,j-1,i) + CZ(co,k-1,j-1,i)) * 0.25;
      }
  if (L->bctype[0] == 7) {   /* metrics.c:690-711 (cgrid == 0) */
    for (int k = gzs + 1; k < gze; k++)
      for (int j = gys + 1; j < gye; j++) {
        V(cx,k,j,-1).x = V(cx,k,j,0).x - V(gs,k,j,-2).x;
        V(cx,k,j,-1).y = V(cx,k,j,0).y;
        V(cx,k,j,-1).z = V(cx,k,j,0).z;
        V(cx,k,j,mx-1).x = V(cx,k,j,mx-2).x + V(gs,k,j,mx+1).x;
        V(cx,k,j,mx-1).y = V(cx,k,j,mx-2).y;
        V(cx,k,j,mx-1).z = V(cx,k,j,mx-2).z;
      }
  }
  /* I-face metrics (metrics.c:713-790) */
  for (int k = lzs; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = 0; i < lxe; i++) {
        if (i == 0 && L->bctype[0] != 7) {
          dxdc = CX(cx,k,j,i+1) - CX(cx,k,j,i); dydc = CY(cx,k,j,i+1) - CY(cx,k,j,i); dzdc = CZ(cx,k,j,i+1) - CZ(cx,k,j,i);
        } else if (i == mx - 2 && L->bctype[0] != 7) {
          dxdc = CX(cx,k,j,i) - CX(cx,k,j,i-1); dydc = CY(cx,k,j,i) - CY(cx,k,j,i-1); dzdc = CZ(cx,k,j,i) - CZ(cx,k,j,i-1);
        } else {
          dxdc = (CX(cx,k,j,i+1) - CX(cx,k,j,i-1)) * 0.5; dydc = (CY(cx,k,j,i+1) - CY(cx,k,j,i-1)) * 0.5;
          dzdc = (CZ(cx,k,j,i+1) - CZ(cx,k,j,i-1)) * 0.5;
        }
        if (j > 1 && j < my - 2) {
          dxde = (CX(cx,k,j+1,i) - CX(cx,k,j-1,i)) * 0.5; dyde = (CY(cx,k,j+1,i) - CY(cx,k,j-1,i)) * 0.5;
          dzde = (CZ(cx,k,j+1,i) - CZ(cx,k,j-1,i)) * 0.5;
        } else if (j == 1) {
          dxde = CX(cx,k,j+1,i) - CX(cx,k,j,i); dyde = CY(cx,k,j+1,i) - CY(cx,k,j,i); dzde = CZ(cx,k,j+1,i) - CZ(cx,k,j,i);
        } else {
          dxde = CX(cx,k,j,i) - CX(cx,k,j-1,i); dyde = CY(cx,k,j,i) - CY(cx,k,j-1,i); dzde = CZ(cx,k,j,i) - CZ(cx,k,j-1,i);
        }
        if (k > 1 && k < mz - 2) {
          dxdz = (CX(cx,k+1,j,i) - CX(cx,k-1,j,i)) * 0.5; dydz = (CY(cx,k+1,j,i) - CY(cx,k-1,j,i)) * 0.5;
          dzdz = (CZ(cx,k+1,j,i) - CZ(cx,k-1,j,i)) * 0.5;
        } else if (k == 1) {
          dxdz = CX(cx,k+1,j,i) - CX(cx,k,j,i); dydz = CY(cx,k+1,j,i) - CY(cx,k,j,i); dzdz = CZ(cx,k+1,j,i) - CZ(cx,k,j,i);
        } else {
          dxdz = CX(cx,k,j,i) - CX(cx,k-1,j,i); dydz = CY(cx,k,j,i) - CY(cx,k-1,j,i); dzdz = CZ(cx,k,j,i) - CZ(cx,k-1,j,i);
        }
        cross_metric(dxde, dyde, dzde, dxdz, dydz, dzdz, &V(&L->icsi, k, j, i));
        cross_metric(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &V(&L->ieta, k, j, i));
        cross_metric(dxdc, dydc, dzdc, dxde, dyde, dzde, &V(&L->izet, k, j, i));
        double det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                     dzdc * (dxde * dydz - dyde * dxdz);
        S(&L->iaj, k, j, i) = 1. / det;
      }

  /* j-face centres (metrics.c:806-819) */
  ofield *cy = &L->centy;
  for (int k = gzs + 1; k < gze; k++)
    for (int j = gys; j < gye; j++)
      for (int i = gxs + 1; i < gxe; i++) {
        if (!(i > 0 && j > -1 && k > 0 && i < IM + 1 && j < JM + 1 && k < KM + 1)) continue;
        V(cy,k,j,i).x = (CX(co,k,j,i) + CX(co,k-1,j,i) + CX(co,k,j,i-1) + CX(co,k-1,j,i-1)) * 0.25;
        V(cy,k,j,i).y = (CY(co,k,j,i) + CY(co,k-1,j,i) + CY(co,k,j,i-1) + CY(co,k-1,j,i-1)) * 0.25;
        V(cy,k,j,i).z = (CZ(co,k,j,i) + CZ(co,k-1,j,i) + CZ(co,k,j,i-1) + CZ(co,k-1,j,i-1)) * 0.25;
      }
  /* J-face metrics (metrics.c:872-942); the i and k differences are
   * one-sided at 1 and m-2 whatever the boundary type */
  for (int k = lzs; k < lze; k++)
    for (int j = 0; j < lye; j++)
      for (int i = lxs; i < lxe; i++) {
        if (i > 1 && i < mx - 2) {
          dxdc = (CX(cy,k,j,i+1) - CX(cy,k,j,i-1)) * 0.5; dydc = (CY(cy,k,j,i+1) - CY(cy,k,j,i-1)) * 0.5;
          dzdc = (CZ(cy,k,j,i+1) - CZ(cy,k,j,i-1)) * 0.5;
        } else if (i == 1) {
          dxdc = CX(cy,k,j,i+1) - CX(cy,k,j,i); dydc = CY(cy,k,j,i+1) - CY(cy,k,j,i); dzdc = CZ(cy,k,j,i+1) - CZ(cy,k,j,i);
        } else {
          dxdc = CX(cy,k,j,i) - CX(cy,k,j,i-1); dydc = CY(cy,k,j,i) - CY(cy,k,j,i-1); dzdc = CZ(cy,k,j,i) - CZ(cy,k,j,i-1);
        }
        if (j == 0) {          /* bctype[2] != 7 always here */
          dxde = CX(cy,k,j+1,i) - CX(cy,k,j,i); dyde = CY(cy,k,j+1,i) - CY(cy,k,j,i); dzde = CZ(cy,k,j+1,i) - CZ(cy,k,j,i);
        } else if (j == my - 2) {
          dxde = CX(cy,k,j,i) - CX(cy,k,j-1,i); dyde = CY(cy,k,j,i) - CY(cy,k,j-1,i); dzde = CZ(cy,k,j,i) - CZ(cy,k,j-1,i);
        } else {
          dxde = (CX(cy,k,j+1,i) - CX(cy,k,j-1,i)) * 0.5; dyde = (CY(cy,k,j+1,i) - CY(cy,k,j-1,i)) * 0.5;
          dzde = (CZ(cy,k,j+1,i) - CZ(cy,k,j-1,i)) * 0.5;
        }
        if (k > 1 && k < mz - 2) {
          dxdz = (CX(cy,k+1,j,i) - CX(cy,k-1,j,i)) * 0.5; dydz = (CY(cy,k+1,j,i) - CY(cy,k-1,j,i)) * 0.5;
          dzdz = (CZ(cy,k+1,j,i) - CZ(cy,k-1,j,i)) * 0.5;
        } else if (k == 1) {
          dxdz = CX(cy,k+1,j,i) - CX(cy,k,j,i); dydz = CY(cy,k+1,j,i) - CY(cy,k,j,i); dzdz = CZ(cy,k+1,j,i) - CZ(cy,k,j,i);
        } else {
          dxdz = CX(cy,k,j,i) - CX(cy,k-1,j,i); dydz = CY(cy,k,j,i) - CY(cy,k-1,j,i); dzdz = CZ(cy,k,j,i) - CZ(cy,k-1,j,i);
        }
        cross_metric(dxde, dyde, dzde, dxdz, dydz, dzdz, &V(&L->jcsi, k, j, i));
        cross_metric(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &V(&L->jeta, k, j, i));
        cross_metric(dxdc, dydc, dzdc, dxde, dyde, dzde, &V(&L->jzet, k, j, i));
        double det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                     dzdc * (dxde * dydz - dyde * dxdz);
        S(&L->jaj, k, j, i) = 1. / det;
      }

  /* k-face centres (metrics.c:956-969) */
  ofield *cz = &L->centz;
  for (int k = gzs; k < gze; k++)
    for (int j = gys + 1; j < gye; j++)
      for (int i = gxs + 1; i < gxe; i++) {
        if (!(i > 0 && j > 0 && k > -1 && i < IM + 1 && j < JM + 1 && k < KM + 1)) continue;
        V(cz,k,j,i).x = (CX(co,k,j,i) + CX(co,k,j-1,i) + CX(co,k,j,i-1) + CX(co,k,j-1,i-1)) * 0.25;
        V(cz,k,j,i).y = (CY(co,k,j,i) + CY(co,k,j-1,i) + CY(co,k,j,i-1) + CY(co,k,j-1,i-1)) * 0.25;
        V(cz,k,j,i).z = (CZ(co,k,j,i) + CZ(co,k,j-1,i) + CZ(co,k,j,i-1) + CZ(co,k,j-1,i-1)) * 0.25;
      }
  if (L->bctype[4] == 7) {   /* metrics.c:996-1017 (cgrid == 0) */
    for (int j = gys + 1; j < gye; j++)
      for (int i = gxs + 1; i < gxe; i++) {
        V(cz,-1,j,i).x = V(cz,0,j,i).x;
        V(cz,-1,j,i).y = V(cz,0,j,i).y;
        V(cz,-1,j,i).z = V(cz,0,j,i).z - V(gs,-2,j,i).z;
        V(cz,mz-1,j,i).x = V(cz,mz-2,j,i).x;
        V(cz,mz-1,j,i).y = V(cz,mz-2,j,i).y;
        V(cz,mz-1,j,i).z = V(cz,mz-2,j,i).z + V(gs,mz+1,j,i).z;
      }
  }
  /* K-face metrics (metrics.c:1019-1090) */
  for (int k = 0; k < lze; k++)
    for (int j = lys; j < lye; j++)
      for (int i = lxs; i < lxe; i++) {
        if (i > 1 && i < mx - 2) {
          dxdc = (CX(cz,k,j,i+1) - CX(cz,k,j,i-1)) * 0.5; dydc = (CY(cz,k,j,i+1) - CY(cz,k,j,i-1)) * 0.5;
          dzdc = (CZ(cz,k,j,i+1) - CZ(cz,k,j,i-1)) * 0.5;
        } else if (i == 1) {
          dxdc = CX(cz,k,j,i+1) - CX(cz,k,j,i); dydc = CY(cz,k,j,i+1) - CY(cz,k,j,i); dzdc = CZ(cz,k,j,i+1) - CZ(cz,k,j,i);
        } else {
          dxdc = CX(cz,k,j,i) - CX(cz,k,j,i-1); dydc = CY(cz,k,j,i) - CY(cz,k,j,i-1); dzdc = CZ(cz,k,j,i) - CZ(cz,k,j,i-1);
        }
        if (j > 1 && j < my - 2) {
          dxde = (CX(cz,k,j+1,i) - CX(cz,k,j-1,i)) * 0.5; dyde = (CY(cz,k,j+1,i) - CY(cz,k,j-1,i)) * 0.5;
          dzde = (CZ(cz,k,j+1,i) - CZ(cz,k,j-1,i)) * 0.5;
        } else if (j == 1) {
          dxde = CX(cz,k,j+1,i) - CX(cz,k,j,i); dyde = CY(cz,k,j+1,i) - CY(cz,k,j,i); dzde = CZ(cz,k,j+1,i) - CZ(cz,k,j,i);
        } else {
          dxde = CX(cz,k,j,i) - CX(cz,k,j-1,i); dyde = CY(cz,k,j,i) - CY(cz,k,j-1,i); dzde = CZ(cz,k,j,i) - CZ(cz,k,j-1,i);
        }
        if (k == 0 && L->bctype[4] != 7) {
          dxdz = CX(cz,k+1,j,i) - CX(cz,k,j,i); dydz = CY(cz,k+1,j,i) - CY(cz,k,j,i); dzdz = CZ(cz,k+1,j,i) - CZ(cz,k,j,i);
        } else if (k == mz - 2 && L->bctype[4] != 7) {
          dxdz = CX(cz,k,j,i) - CX(cz,k-1,j,i); dydz = CY(cz,k,j,i) - CY(cz,k-1,j,i); dzdz = CZ(cz,k,j,i) - CZ(cz,k-1,j,i);
        } else {
          dxdz = (CX(cz,k+1,j,i) - CX(cz,k-1,j,i)) * 0.5; dydz = (CY(cz,k+1,j,i) - CY(cz,k-1,j,i)) * 0.5;
          dzdz = (CZ(cz,k+1,j,i) - CZ(cz,k-1,j,i)) * 0.5;
        }
        cross_metric(dxde, dyde, dzde, dxdz, dydz, dzdz, &V(&L->kcsi, k, j, i));
        cross_metric(dxdz, dydz, dzdz, dxdc, dydc, dzdc, &V(&L->keta, k, j, i));
        cross_metric(dxdc, dydc, dzdc, dxde, dyde, dzde, &V(&L->kzet, k, j, i));
        double det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                     dzdc * (dxde * dydz - dyde * dxdz);
        S(&L->kaj, k, j, i) = 1. / det;
      }

  /* global -> local for every metric (metrics.c:1151-1203) */
  ofield *all[] = {csi, eta, zet, aj, &L->icsi, &L->ieta, &L->izet, &L->iaj,
                   &L->jcsi, &L->jeta, &L->jzet, &L->jaj, &L->kcsi, &L->keta, &L->kzet, &L->kaj, cent};
  for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); q++) o_g2l(all[q], L->per);

  /* Update_Metrics_PBC (implicitsolver.c:2174-2573): in a periodic axis the
   * local arrays take [0]<-[-2], [-1]<-[-3], [m-1]<-[m+1], [m]<-[m+2] */
  ofield *pbc[] = {csi, eta, zet, &L->icsi, &L->ieta, &L->izet, &L->jcsi, &L->jeta, &L->jzet,
                   &L->kcsi, &L->keta, &L->kzet, aj, &L->iaj, &L->jaj, &L->kaj};
  const int npbc = sizeof(pbc) / sizeof(pbc[0]);
  if (L->bctype[0] == 7) {
    for (int q = 0; q < npbc; q++) {
      ofield *f = pbc[q]; int d = f->dof;
      for (int k = gzs; k < gze; k++)
        for (int j = gys; j < gye; j++) {
          for (int c = 0; c < d; c++) {
            f->a[oix(f,0,j,k)*d+c] = f->a[oix(f,-2,j,k)*d+c];
            f->a[oix(f,-1,j,k)*d+c] = f->a[oix(f,-3,j,k)*d+c];
          }
        }
      for (int k = gzs; k < gze; k++)
        for (int j = gys; j < gye; j++)
          for (int c = 0; c < d; c++) {
            f->a[oix(f,mx-1,j,k)*d+c] = f->a[oix(f,mx+1,j,k)*d+c];
            f->a[oix(f,mx,j,k)*d+c] = f->a[oix(f,mx+2,j,k)*d+c];
          }
    }
  }
  if (L->bctype[4] == 7) {
    for (int q = 0; q < npbc; q++) {
      ofield *f = pbc[q]; int d = f->dof;
      for (int j = gys; j < gye; j++)
        for (int i = gxs; i < gxe; i++)
          for (int c = 0; c < d; c++) {
            f->a[oix(f,i,j,0)*d+c] = f->a[oix(f,i,j,-2)*d+c];
            f->a[oix(f,i,j,-1)*d+c] = f->a[oix(f,i,j,-3)*d+c];
          }
      for (int j = gys; j < gye; j++)
        for (int i = gxs; i < gxe; i++)
          for (int c = 0; c < d; c++) {
            f->a[oix(f,i,j,mz-1)*d+c] = f->a[oix(f,i,j,mz+1)*d+c];
            f->a[oix(f,i,j,mz)*d+c] = f->a[oix(f,i,j,mz+2)*d+c];
          }
    }
  }
}

void o_form_metrics_all(octx *c) {
  for (int l = c->nlev - 1; l >= 0; l--) form_metrics_level(&c->lev[l]);
}

/* ------------------------------------------------------------------ */
/* nvert                                                                */
/* ------------------------------------------------------------------ */
void o_set_nvert(octx *c, const double *nv) {
  olevel *F = &c->lev[c->nlev - 1];
  o_put_field(&F->nvert, nv);
  o_g2l(&F->nvert, F->per);
}

/* MyNvertRestriction (poisson.c:1662-1773) */
static void nvert_restrict(olevel *H, olevel *Lc) {
  const int mx = Lc->mx, my = Lc->my, mz = Lc->mz;
  int ia = Lc->isc ? 0 : 1, ja = Lc->jsc ? 0 : 1, ka = Lc->ksc ? 0 : 1;
  ofield *nh = &H->nvert, *nc = &Lc->nvert;
  of_zero(nc);
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        int ih = Lc->isc ? i : 2 * i, jh = Lc->jsc ? j : 2 * j, kh = Lc->ksc ? k : 2 * k;
        double prod = S(nh,kh,jh,ih) * S(nh,kh,jh,ih-ia) * S(nh,kh,jh-ja,ih) * S(nh,kh-ka,jh,ih) *
                      S(nh,kh,jh-ja,ih-ia) * S(nh,kh-ka,jh,ih-ia) * S(nh,kh-ka,jh-ja,ih) *
                      S(nh,kh-ka,jh-ja,ih-ia);
        if (prod > 0.1) S(nc,k,j,i) = fmax(1., S(nc,k,j,i));
      }
  o_g2l(nc, Lc->per);
  /* second pass reads the local copy, writes the global one */
  ofield tmp; of_alloc(&tmp, mx, my, mz, 1); of_copy(&tmp, nc);
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        if (S(&tmp,k,j,i) < 0.1) {
          if (S(&tmp,k,j,i+1) + S(&tmp,k,j,i-1) > 1.1 && S(&tmp,k,j+1,i) + S(&tmp,k,j-1,i) > 1.1 &&
              S(&tmp,k+1,j,i) + S(&tmp,k-1,j,i) > 1.1)
            S(nc,k,j,i) = 1.;
        }
      }
  of_free(&tmp);
  o_g2l(nc, Lc->per);
}

void o_nvert_restrict_all(octx *c) {
  for (int l = c->nlev - 1; l > 0; l--) nvert_restrict(&c->lev[l], &c->lev[l - 1]);
}
