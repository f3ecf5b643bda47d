/*
 * cv_oracle_momentum.c -- TEST INFRASTRUCTURE (CPU oracle): momentum right-
 * hand side and the step driver.
 *
 *   o_contra2cart       Contra2Cart          rhs.c:25-144
 *   convection          Convection (QUICK)   rhs.c:1463-1945
 *   viscous             Viscous (laminar)    rhs.c:1947-2565
 *   o_formfunction1     FormFunction1        rhs.c:2567-3351
 *   o_inflowflux        InflowFlux, analytic profiles  bcs.c:99-246
 *   o_outflowflux       OutflowFlux          bcs.c:403-469
 *   o_formbcs           FormBCS, bctype {1,3,4,5,6,7}  bcs.c:4436-5888
 *   o_ghostnodevelocity GhostNodeVelocity    bcs.c:3645-4435 (subset)
 *   o_rungekutta        RungeKutta           implicitsolver.c:32-119
 *   o_flow_step         Flow_Solver          solvers.c:944-1156 (+ the level
 *                       rotation Ucont_o <- Ucont of main.c:1123-1145)
 *
 * Parity pins (closed forms; tests/test_oracle_momentum.py unless noted):
 *  - Convection: QUICK face values of a cubic profile carried by a flux that
 *    changes sign (exact for the quadratic part, the Lagrange remainder
 *    (3/8) e h^3 sign(F) for the cubic part);
 *  - Viscous: a quadratic velocity on a sheared lattice gives the viscous
 *    right-hand side (nu1 - nu2) A.(lap u + grad div u) exactly; Poiseuille
 *    flow on a uniform grid (test_oracle_pins.py);
 *  - FormFunction1: -(face area).grad p for a linear p on uniform and
 *    sheared lattices; a divergence-free shear wave on a lattice whose j
 *    face areas vary along j (the averaged basis of rct); free-stream
 *    preservation on the warped channel;
 *  - RungeKutta: a viscous shear-wave mode is multiplied by
 *    1 + s + s^2/2 + s^3/6 + s^4/24;
 *  - Contra2Cart: constant flow round trip on a warped grid with a body;
 *    CalcRHS/ImpRK: BDF1/BDF2 exactness, the boundary zero set, the stopping
 *    test (test_oracle_pins.py).
 *  Boundaries (test_oracle_pins.py): the inlet profiles against their
 *  closed forms (the duct series solves u_xx + u_yy = -2 with u = 0 on the
 *  duct walls), the inlet flux = uin * face area on a uniform grid, the
 *  outlet correction's mass balance (fluid outflow = FluxInSum) and a
 *  uniform stream through an inlet/outlet box left unchanged by FormBCS.
 *  tests/test_oracle_mutations.py plants a mistake in each of these
 *  routines and checks that a pin fails.  The stale dpdc/dpde/dpdz carry of
 *  FormFunction1 (rhs.c:2803-2861) is reproduced in loop order; its first
 *  read is of an uninitialised variable in the reference and is 0 here;
 *  no closed form reaches it (parity unpinned).
 *
 * Deliberate deviations (the GPU library makes the same ones):
 *  - no-slip walls (bctype 1): the reference never resets the wall-normal
 *    face flux, so FormFunction1's interior loop drives the flux through the
 *    upper wall face (j=my-2 etc.) while the lower one keeps its initial
 *    value, which makes the Poisson right-hand side incompatible.  After each
 *    RK stage the wall-normal flux of bctype-1 faces is set to the wall
 *    value 0, as FormBCS does for the slip wall (bcs.c:4745-4853).
 *  - GhostNodeVelocity's two periodic copies that read ".x" from k+2
 *    (bcs.c:4099, 4159) are taken as full copies from the wrapped plane.
 */
#include "cv_oracle_int.h"
#include <stdio.h>
#include <time.h>

int o_xmode(double hi_pair, double lo_pair, int at_hi, int at_lo, int per);

static double now_s(void) {
  struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* Contra2Cart (rhs.c:25-144)                                           */
/* ------------------------------------------------------------------ */
void o_contra2cart(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const ofield *csi = &L->csi, *eta = &L->eta, *zet = &L->zet, *nv = &L->nvert;
  const ofield *uc = &c->ucont;
  ofield *ucat = &c->ucat;
  const double innerblank = 7.;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        double n = S(nv, k, j, i);
        if (!(n < 0.1 || n > innerblank)) continue;
        double mat[3][3], q[3];
        mat[0][0] = 0.5 * (V(csi,k,j,i-1).x + V(csi,k,j,i).x);
        mat[0][1] = 0.5 * (V(csi,k,j,i-1).y + V(csi,k,j,i).y);
        mat[0][2] = 0.5 * (V(csi,k,j,i-1).z + V(csi,k,j,i).z);
        mat[1][0] = 0.5 * (V(eta,k,j-1,i).x + V(eta,k,j,i).x);
        mat[1][1] = 0.5 * (V(eta,k,j-1,i).y + V(eta,k,j,i).y);
        mat[1][2] = 0.5 * (V(eta,k,j-1,i).z + V(eta,k,j,i).z);
        mat[2][0] = 0.5 * (V(zet,k-1,j,i).x + V(zet,k,j,i).x);
        mat[2][1] = 0.5 * (V(zet,k-1,j,i).y + V(zet,k,j,i).y);
        mat[2][2] = 0.5 * (V(zet,k-1,j,i).z + V(zet,k,j,i).z);
        q[0] = 0.5 * (V(uc,k,j,i-1).x + V(uc,k,j,i).x);
        q[1] = 0.5 * (V(uc,k,j-1,i).y + V(uc,k,j,i).y);
        q[2] = 0.5 * (V(uc,k-1,j,i).z + V(uc,k,j,i).z);
        double det = mat[0][0] * (mat[1][1] * mat[2][2] - mat[1][2] * mat[2][1]) -
                     mat[0][1] * (mat[1][0] * mat[2][2] - mat[1][2] * mat[2][0]) +
                     mat[0][2] * (mat[1][0] * mat[2][1] - mat[1][1] * mat[2][0]);
        double det0 = q[0] * (mat[1][1] * mat[2][2] - mat[1][2] * mat[2][1]) -
                      q[1] * (mat[0][1] * mat[2][2] - mat[0][2] * mat[2][1]) +
                      q[2] * (mat[0][1] * mat[1][2] - mat[0][2] * mat[1][1]);
        double det1 = -q[0] * (mat[1][0] * mat[2][2] - mat[1][2] * mat[2][0]) +
                      q[1] * (mat[0][0] * mat[2][2] - mat[0][2] * mat[2][0]) -
                      q[2] * (mat[0][0] * mat[1][2] - mat[0][2] * mat[1][0]);
        double det2 = q[0] * (mat[1][0] * mat[2][1] - mat[1][1] * mat[2][0]) -
                      q[1] * (mat[0][0] * mat[2][1] - mat[0][1] * mat[2][0]) +
                      q[2] * (mat[0][0] * mat[1][1] - mat[0][1] * mat[1][0]);
        V(ucat,k,j,i).x = det0 / det;
        V(ucat,k,j,i).y = det1 / det;
        V(ucat,k,j,i).z = det2 / det;
      }
  o_g2l(ucat, L->per);
}

/* ------------------------------------------------------------------ */
/* ghost planes of Ucat/P/Nvert (FormBCS and GhostNodeVelocity)         */
/* ------------------------------------------------------------------ */
/* wall ghost: ucat_g = 2*ubcs - ucat_inner (bcs.c:5252-5314);
 * full=1 covers the full tangential ranges (GhostNodeVelocity, bcs.c:4015-4081),
 * full=0 the interior ranges (FormBCS). */
static void wall_ghosts(octx *c, int full) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  ofield *u = &c->ucat; const ofield *ub = &c->ubcs;
  int a0 = full ? 0 : 1, ax = full ? mx : mx - 1, ay = full ? my : my - 1, az = full ? mz : mz - 1;
#define WG(kk, jj, ii, kn, jn, in)                                  \
  do {                                                              \
    V(u,kk,jj,ii).x = 2 * V(ub,kk,jj,ii).x - V(u,kn,jn,in).x;        \
    V(u,kk,jj,ii).y = 2 * V(ub,kk,jj,ii).y - V(u,kn,jn,in).y;        \
    V(u,kk,jj,ii).z = 2 * V(ub,kk,jj,ii).z - V(u,kn,jn,in).z;        \
  } while (0)
  if (L->bctype[0] != 7) {
    for (int k = a0; k < az; k++) for (int j = a0; j < ay; j++) WG(k, j, 0, k, j, 1);
    for (int k = a0; k < az; k++) for (int j = a0; j < ay; j++) WG(k, j, mx - 1, k, j, mx - 2);
  }
  if (L->bctype[2] != 7) {
    for (int k = a0; k < az; k++) for (int i = a0; i < ax; i++) WG(k, 0, i, k, 1, i);
    for (int k = a0; k < az; k++) for (int i = a0; i < ax; i++) WG(k, my - 1, i, k, my - 2, i);
  }
  if (L->bctype[4] != 7) {
    for (int j = a0; j < ay; j++) for (int i = a0; i < ax; i++) WG(0, j, i, 1, j, i);
    for (int j = a0; j < ay; j++) for (int i = a0; i < ax; i++) WG(mz - 1, j, i, mz - 2, j, i);
  }
#undef WG
}

/* periodic plane copies restricted to the interior of the other axes
 * (bcs.c:5325-5431) */
static void periodic_planes(octx *c, int with_p) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz, IM = L->IM, JM = L->JM, KM = L->KM;
  if (!(L->bctype[0] == 7 || L->bctype[4] == 7)) return;
  ofield *u = &c->ucat, *p = &c->p, *nv = &L->nvert;
  o_g2l(u, L->per); if (with_p) { o_g2l(p, L->per); }
  /* the copies read the local (pre-copy) arrays */
  ofield lu, lp, ln;
  of_alloc(&lu, mx, my, mz, 3); of_copy(&lu, u);
  of_alloc(&lp, mx, my, mz, 1); of_copy(&lp, p);
  of_alloc(&ln, mx, my, mz, 1); of_copy(&ln, nv);
  if (L->bctype[0] == 7)
    for (int k = 0; k < mz; k++) for (int j = 0; j < my; j++)
      if (j > 0 && k > 0 && j < JM && k < KM) {
        V(u,k,j,0) = V(&lu,k,j,-2);
        if (with_p) { S(p,k,j,0) = S(&lp,k,j,-2); S(nv,k,j,0) = S(&ln,k,j,-2); }
      }
  if (L->bctype[4] == 7)
    for (int j = 0; j < my; j++) for (int i = 0; i < mx; i++)
      if (i > 0 && j > 0 && i < IM && j < JM) {
        V(u,0,j,i) = V(&lu,-2,j,i);
        if (with_p) { S(p,0,j,i) = S(&lp,-2,j,i); S(nv,0,j,i) = S(&ln,-2,j,i); }
      }
  if (L->bctype[0] == 7)
    for (int k = 0; k < mz; k++) for (int j = 0; j < my; j++)
      if (j > 0 && k > 0 && j < JM && k < KM) {
        V(u,k,j,mx-1) = V(&lu,k,j,mx+1);
        if (with_p) { S(p,k,j,mx-1) = S(&lp,k,j,mx+1); S(nv,k,j,mx-1) = S(&ln,k,j,mx+1); }
      }
  if (L->bctype[4] == 7)
    for (int j = 0; j < my; j++) for (int i = 0; i < mx; i++)
      if (i > 0 && j > 0 && i < IM && j < JM) {
        V(u,mz-1,j,i) = V(&lu,mz+1,j,i);
        if (with_p) { S(p,mz-1,j,i) = S(&lp,mz+1,j,i); S(nv,mz-1,j,i) = S(&ln,mz+1,j,i); }
      }
  of_free(&lu); of_free(&lp); of_free(&ln);
  o_g2l(u, L->per); o_g2l(p, L->per); o_g2l(nv, L->per);
}

/* edge values: average of the two adjacent ghost-plane values
 * ("0 velocity on the corner point", bcs.c:5440-5566), in the reference's
 * sequential order */
static void edge_average(octx *c, int with_p) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  ofield *u = &c->ucat, *p = &c->p;
#define EA(k0, j0, i0, ka, ja, ia, kb, jb, ib)                                   \
  do {                                                                           \
    V(u,k0,j0,i0).x = 0.5 * (V(u,ka,ja,ia).x + V(u,kb,jb,ib).x);                  \
    V(u,k0,j0,i0).y = 0.5 * (V(u,ka,ja,ia).y + V(u,kb,jb,ib).y);                  \
    V(u,k0,j0,i0).z = 0.5 * (V(u,ka,ja,ia).z + V(u,kb,jb,ib).z);                  \
    if (with_p) S(p,k0,j0,i0) = 0.5 * (S(p,ka,ja,ia) + S(p,kb,jb,ib));           \
  } while (0)
  for (int side = 0; side < 2; side++) {
    int k = side ? mz - 1 : 0, kn = side ? mz - 2 : 1;
    for (int j = 0; j < my; j++) EA(k, j, 0, kn, j, 0, k, j, 1);
    for (int j = 0; j < my; j++) EA(k, j, mx - 1, kn, j, mx - 1, k, j, mx - 2);
    for (int i = 0; i < mx; i++) EA(k, 0, i, kn, 0, i, k, 1, i);
    for (int i = 0; i < mx; i++) EA(k, my - 1, i, kn, my - 1, i, k, my - 2, i);
  }
  for (int side = 0; side < 2; side++) {
    int j = side ? my - 1 : 0, jn = side ? my - 2 : 1;
    for (int k = 0; k < mz; k++) EA(k, j, 0, k, jn, 0, k, j, 1);
    for (int k = 0; k < mz; k++) EA(k, j, mx - 1, k, jn, mx - 1, k, j, mx - 2);
  }
#undef EA
  o_g2l(u, L->per); o_g2l(p, L->per);
}

/* periodic full-plane copies (bcs.c:5567-5731); the reference copies only
 * .x on the i=mx-1 plane here (bcs.c:5613) */
static void periodic_corners(octx *c, int with_p, int x_only_upper) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  if (!(L->bctype[0] == 7 || L->bctype[4] == 7)) return;
  ofield *u = &c->ucat, *p = &c->p, *nv = &L->nvert;
  if (L->bctype[0] == 7) {
    for (int k = 0; k < mz; k++) for (int j = 0; j < my; j++) {
      V(u,k,j,0) = V(u,k,j,-2);
      if (with_p) { S(p,k,j,0) = S(p,k,j,-2); S(nv,k,j,0) = S(nv,k,j,-2); }
    }
    for (int k = 0; k < mz; k++) for (int j = 0; j < my; j++) {
      if (x_only_upper) V(u,k,j,mx-1).x = V(u,k,j,mx+1).x;
      else V(u,k,j,mx-1) = V(u,k,j,mx+1);
      if (with_p) { S(p,k,j,mx-1) = S(p,k,j,mx+1); S(nv,k,j,mx-1) = S(nv,k,j,mx+1); }
    }
    o_g2l(u, L->per); o_g2l(p, L->per); o_g2l(nv, L->per);
  }
  if (L->bctype[4] == 7) {
    for (int j = 0; j < my; j++) for (int i = 0; i < mx; i++) {
      V(u,0,j,i) = V(u,-2,j,i);
      if (with_p) { S(p,0,j,i) = S(p,-2,j,i); S(nv,0,j,i) = S(nv,-2,j,i); }
    }
    for (int j = 0; j < my; j++) for (int i = 0; i < mx; i++) {
      V(u,mz-1,j,i) = V(u,mz+1,j,i);
      if (with_p) { S(p,mz-1,j,i) = S(p,mz+1,j,i); S(nv,mz-1,j,i) = S(nv,mz+1,j,i); }
    }
    o_g2l(u, L->per); o_g2l(p, L->per); o_g2l(nv, L->per);
  }
}

/* zero the wall-normal flux of bctype-1 faces, each face on its own
 * (deviation, see header) */
static void wall_flux_zero(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  ofield *u = &c->ucont;
  for (int k = 0; k < mz; k++) for (int j = 0; j < my; j++) {
    if (L->bctype[0] == 1) V(u,k,j,0).x = 0.;
    if (L->bctype[1] == 1) V(u,k,j,mx-2).x = 0.;
  }
  for (int k = 0; k < mz; k++) for (int i = 0; i < mx; i++) {
    if (L->bctype[2] == 1) V(u,k,0,i).y = 0.;
    if (L->bctype[3] == 1) V(u,k,my-2,i).y = 0.;
  }
  for (int j = 0; j < my; j++) for (int i = 0; i < mx; i++) {
    if (L->bctype[4] == 1) V(u,0,j,i).z = 0.;
    if (L->bctype[5] == 1) V(u,mz-2,j,i).z = 0.;
  }
  o_g2l(u, L->per);
}

/* ------------------------------------------------------------------ */
/* InflowFlux (bcs.c:99-246) for an inlet on the k- face (bctype 5)      */
/* ------------------------------------------------------------------ */
void o_set_inlet(octx *c, int profile, double x_c, double y_c, double z_c) {
  c->inletprofile = profile; c->cm_c[0] = x_c; c->cm_c[1] = y_c; c->cm_c[2] = z_c;
}

/* the analytic inlet profiles (bcs.c:158-196); 0, 3, 5, 6 and 8 read an
 * inlet file (InletInterpolation) and are not restated */
int o_inlet_profile_ok(int profile) {
  return profile == 1 || profile == -1 || profile == 2 || profile == 4 || profile == 7 || profile == 10 ||
         profile == 11;
}

double o_inlet_uin(int profile, double xc, double yc, double zc) {
  const double H = 4.1, Umax = 1.5;
  const double r = sqrt(zc * zc + yc * yc);
  double uin = 0.;
  if (profile == 1) uin = 2.;
  else if (profile == -1) uin = -1.;
  else if (profile == 2) uin = 4. * Umax * yc * (H - yc) / (H * H);
  else if (profile == 4) {           /* fully developed pipe flow */
    uin = 2. * (1. - 4. * r * r);
    if (r > 0.5) uin = 0.;
  } else if (profile == 7) uin = 1.5 * (1. - 4. * yc * yc);
  else if (profile == 10) {          /* fully developed flow in a 2 x 2 duct, 20 terms */
    const double pi = 3.14159265358979323846264338327950288;
    const double y2 = (yc - 1.5) * 2., x2 = xc - 0.5, A = 1.;
    uin = 1 - y2 * y2;
    for (int n = 0; n < 20; n++)
      uin -= 4. * pow(2. / pi, 3) * pow(-1., n) / pow(2 * n + 1., 3.) * cosh((2 * n + 1) * pi * x2 / 2.) *
             cos((2. * n + 1) * pi * y2 / 2.) / cosh((2 * n + 1) * pi * A / 2.);
  } else if (profile == 11) uin = 0.185;
  return uin;
}

/* FluxInSum is a global that only the inlet branch writes (main.c:9) */
void o_inflowflux(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  if (L->bctype[4] != 5) return;
  const int mx = L->mx, my = L->my;
  const ofield *co = &L->coor, *zet = &L->zet, *nv = &L->nvert;
  ofield *uc = &c->ucont, *ub = &c->ubcs;
  const int k = 0;
  for (int j = 1; j < my - 1; j++)
    for (int i = 1; i < mx - 1; i++) {
      const double xc = (V(co,k,j,i).x + V(co,k,j-1,i).x + V(co,k,j,i-1).x + V(co,k,j-1,i-1).x) * 0.25 - c->cm_c[0];
      const double yc = (V(co,k,j,i).y + V(co,k,j-1,i).y + V(co,k,j,i-1).y + V(co,k,j-1,i-1).y) * 0.25 - c->cm_c[1];
      const double zc = (V(co,k,j,i).z + V(co,k,j-1,i).z + V(co,k,j,i-1).z + V(co,k,j-1,i-1).z) * 0.25 - c->cm_c[2];
      const double uin = o_inlet_uin(c->inletprofile, xc, yc, zc);
      if (S(nv,k+1,j,i) < 0.1) {
        const ocmp z = V(zet,k,j,i);
        const double a = sqrt(z.z * z.z + z.y * z.y + z.x * z.x);
        V(ub,k,j,i).y = uin * z.y / a;
        V(ub,k,j,i).z = uin * z.z / a;
        V(ub,k,j,i).x = uin * z.x / a;
        V(uc,k,j,i).z = uin * a;
      }
    }
  double flux = 0.;
  for (int j = 1; j < my - 1; j++)
    for (int i = 1; i < mx - 1; i++)
      if (S(nv,k+1,j,i) < 0.1) flux += V(uc,k,j,i).z;
  c->flux_in_sum = flux;
  o_g2l(uc, L->per);
}

/* OutflowFlux (bcs.c:403-469): the flux through the k+ outlet face,
 * fluid or not.  FormBCS overwrites the FluxOutSum it leaves before any
 * use, so it is a diagnostic only. */
void o_outflowflux(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  if (L->bctype[5] != 4) return;
  const int mx = L->mx, my = L->my, k = L->mz - 2;
  double flux = 0.;
  for (int j = 1; j < my - 1; j++)
    for (int i = 1; i < mx - 1; i++) flux += V(&c->ucont,k,j,i).z;
  c->flux_out_diag = flux;
}

/* point (s, t) of face f (0 i-, 1 i+, 2 j-, 3 j+, 4 k-, 5 k+), s and t the
 * outer and inner indices of the reference's face loops (i faces k, j; j
 * faces k, i; k faces j, i): the ghost cell g, the adjacent cell a and the
 * storage point b of the boundary face (0 on lower faces, m-2 on upper
 * ones), each as (i, j, k) */
static void face_pt(const olevel *L, int f, int s, int t, int g[3], int a[3], int b[3]) {
  const int d = f / 2, up = f % 2, m = d == 0 ? L->mx : (d == 1 ? L->my : L->mz);
  int p[3] = {0, 0, 0};
  if (d == 0) { p[2] = s; p[1] = t; }
  else if (d == 1) { p[2] = s; p[0] = t; }
  else { p[1] = s; p[0] = t; }
  for (int q = 0; q < 3; q++) g[q] = a[q] = b[q] = p[q];
  g[d] = up ? m - 1 : 0; a[d] = up ? m - 2 : 1; b[d] = up ? m - 2 : 0;
}
/* extents of the outer and inner loop axes of face f */
static void face_ext(const olevel *L, int f, int *ms, int *mt) {
  const int d = f / 2;
  *ms = d == 2 ? L->my : L->mz;
  *mt = d == 0 ? L->my : L->mx;
}
static const ofield *face_metric(const olevel *L, int f) {
  return f / 2 == 0 ? &L->csi : (f / 2 == 1 ? &L->eta : &L->zet);
}
#define VP(f, p) V(f, (p)[2], (p)[1], (p)[0])
#define CMP(v, d) (((double *)&(v))[d])

/* slip walls, bctype 3 (FormBCS bcs.c:4729-4853, GhostNodeVelocity
 * bcs.c:3833-3960): Ubcs = the adjacent Ucat less its component along the
 * face normal; FormBCS also zeroes the face's normal flux */
static void slip_faces(octx *c, int formbcs) {
  olevel *L = &c->lev[c->nlev - 1];
  for (int f = 0; f < 6; f++) {
    if (L->bctype[f] != 3) continue;
    const int d = f / 2;
    const ofield *met = face_metric(L, f);
    int ms, mt;
    face_ext(L, f, &ms, &mt);
    for (int s = 1; s < ms - 1; s++)
      for (int t = 1; t < mt - 1; t++) {
        int g[3], a[3], b[3];
        face_pt(L, f, s, t, g, a, b);
        const ocmp m = VP(met, b), u = VP(&c->ucat, a);
        const double A = sqrt(m.z * m.z + m.y * m.y + m.x * m.x);
        const double nx = m.x / A, ny = m.y / A, nz = m.z / A;
        const double Un = u.x * nx + u.y * ny + u.z * nz;
        ocmp *ub = &VP(&c->ubcs, g);
        ub->x = u.x - Un * nx;
        ub->y = u.y - Un * ny;
        ub->z = u.z - Un * nz;
        if (formbcs) CMP(VP(&c->ucont, b), d) = 0.;
      }
  }
}

/* far field, bctype 6 (FormBCS bcs.c:4500-4724): Ubcs = the adjacent Ucat
 * and the face flux its normal component times the matching metric
 * component; when an upper face is far field, a correction proportional to
 * each face's |flux| balances the far-field inflow (lower faces) and
 * outflow (upper faces).  FarFluxInSum feeds the outlet section. */
static void farfield(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  double fin = 0., fout = 0., fabs_sum = 0., ain = 0., aout = 0.;
  for (int f = 0; f < 6; f++) {
    if (L->bctype[f] != 6) continue;
    const int d = f / 2, up = f % 2;
    const ofield *met = face_metric(L, f);
    int ms, mt;
    face_ext(L, f, &ms, &mt);
    for (int s = 1; s < ms - 1; s++)
      for (int t = 1; t < mt - 1; t++) {
        int g[3], a[3], b[3];
        face_pt(L, f, s, t, g, a, b);
        const ocmp u = VP(&c->ucat, a);
        VP(&c->ubcs, g) = u;
        const double mc = CMP(VP(met, b), d);
        const double q = CMP(u, d) * mc;
        CMP(VP(&c->ucont, b), d) = q;
        if (up) { fout += q; aout += mc; }
        else { fin += q; ain += mc; }
        fabs_sum += fabs(q);
      }
  }
  (void)ain; (void)aout;   /* the far-field areas only enter the reference's log line */
  c->far_flux_in = fin; c->far_flux_out = fout;
  double ratio = 0.;
  if (L->bctype[5] == 6 || L->bctype[3] == 6 || L->bctype[1] == 6) {
    ratio = (fin - fout) / fabs_sum;
    if (fabs(fabs_sum) < 1.e-10) ratio = 0.;
  }
  const int order[6] = {5, 3, 1, 0, 2, 4};
  for (int o = 0; o < 6; o++) {
    const int f = order[o];
    if (L->bctype[f] != 6) continue;
    const int d = f / 2, up = f % 2;
    const ofield *met = face_metric(L, f);
    int ms, mt;
    face_ext(L, f, &ms, &mt);
    for (int s = 1; s < ms - 1; s++)
      for (int t = 1; t < mt - 1; t++) {
        int g[3], a[3], b[3];
        face_pt(L, f, s, t, g, a, b);
        double *q = &CMP(VP(&c->ucont, b), d);
        if (up) *q += ratio * fabs(*q);
        else *q -= ratio * fabs(*q);
        CMP(VP(&c->ubcs, g), d) = *q / CMP(VP(met, b), d);
      }
  }
}
#undef VP
#undef CMP

/* FormBCS's outlet section for bctype 4 on the k+ face (bcs.c:4914-5010):
 * Ubcs = Ucat of the last cell, and the outlet flux is that velocity's flux
 * plus a uniform correction that makes the fluid outflow equal FluxInSum */
static void outlet_kplus(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  if (L->bctype[5] != 4) return;
  const int mx = L->mx, my = L->my, k = L->mz - 1;
  const ofield *zet = &L->zet, *nv = &L->nvert, *ucat = &c->ucat;
  ofield *uc = &c->ucont, *ub = &c->ubcs;
  double fout = 0., area = 0.;
  for (int j = 1; j < my - 1; j++)
    for (int i = 1; i < mx - 1; i++)
      if (S(nv,k-1,j,i) < 0.1) {
        const ocmp u = V(ucat,k-1,j,i), z = V(zet,k-1,j,i);
        fout += u.x * z.x + u.y * z.y + u.z * z.z;
        area += sqrt(z.x * z.x + z.y * z.y + z.z * z.z);
      }
  /* FluxIn = FluxInSum + FarFluxInSum + FluxIntpSum, the last 0 here */
  const double ratio = (c->flux_in_sum + c->far_flux_in - fout) / area;
  double fsum = 0.;
  for (int j = 1; j < my - 1; j++)
    for (int i = 1; i < mx - 1; i++)
      if (S(nv,k-1,j,i) < 0.1) {
        V(ub,k,j,i) = V(ucat,k-1,j,i);
        const ocmp b = V(ub,k,j,i), z = V(zet,k-1,j,i);
        V(uc,k-1,j,i).z = (b.x * z.x + b.y * z.y + b.z * z.z) + ratio * sqrt(z.x * z.x + z.y * z.y + z.z * z.z);
        fsum += V(uc,k-1,j,i).z;
      }
  c->flux_out_sum = fsum;
}

/* FormBCS for bctype {1,3,4,5,6,7} (bcs.c:4436-5888): three passes of
 * far field -> slip -> outlet correction -> Contra2Cart -> wall ghosts ->
 * periodic planes -> edge averages -> periodic full planes. */
void o_formbcs(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  for (int ttemp = 0; ttemp < c->formbcs_passes; ttemp++) {
    farfield(c);
    slip_faces(c, 1);
    outlet_kplus(c);
    o_g2l(&c->ucont, L->per);
    o_contra2cart(c);
    wall_ghosts(c, 0);
    periodic_planes(c, 1);
    edge_average(c, 1);
    periodic_corners(c, 1, 1);
  }
  o_g2l(&c->p, L->per); o_g2l(&c->ucat, L->per); o_g2l(&c->ucont, L->per);
}

/* test hook: the number of FormBCS passes (the reference makes 3) */
void o_set_formbcs_passes(octx *c, int n) { c->formbcs_passes = n; }

/* GhostNodeVelocity (bcs.c:3645-4435) for bctype {1,3,4,5,6,7}: slip
 * faces take Ubcs from the current Ucat, the others keep the Ubcs the last
 * InflowFlux / FormBCS left */
void o_ghostnodevelocity(octx *c) {
  slip_faces(c, 0);
  wall_ghosts(c, 1);
  periodic_planes(c, 0);
  edge_average(c, 0);
  periodic_corners(c, 0, 0);
}

/* ------------------------------------------------------------------ */
/* ibm_interpolation_advanced (ibm.c:3012-3254), Add_dUndt = 1, without  */
/* wall functions: six sweeps of the IB-node reconstruction, the fluxes  */
/* of the faces touching an IB node, and Contra2Cart                      */
/* ------------------------------------------------------------------ */
void o_set_ib_nodes(octx *c, int n, const int *idx12, const double *val12, int mode) {
  free(c->ib_idx); free(c->ib_val);
  c->ib_idx = NULL; c->ib_val = NULL; c->ib_n = 0; c->ib_mode = mode;
  if (n <= 0) return;
  c->ib_idx = (int *)malloc((size_t)12 * n * sizeof(int));
  c->ib_val = (double *)malloc((size_t)12 * n * sizeof(double));
  memcpy(c->ib_idx, idx12, (size_t)12 * n * sizeof(int));
  memcpy(c->ib_val, val12, (size_t)12 * n * sizeof(double));
  c->ib_n = n;
}

/* noslip (wallfunction.c:12-27): linear in the distance from the wall */
static void ib_noslip(double sc, double sb, ocmp Ua, ocmp Uc, ocmp *Ub) {
  const double u_c = Uc.x - Ua.x, v_c = Uc.y - Ua.y, w_c = Uc.z - Ua.z;
  Ub->x = sb / sc * u_c;
  Ub->y = sb / sc * v_c;
  Ub->z = sb / sc * w_c;
  Ub->x += Ua.x;
  Ub->y += Ua.y;
  Ub->z += Ua.z;
}

/* freeslip (wallfunction.c:29-46): the normal component linear in the
 * distance, the tangential one that of the interception point */
static void ib_freeslip(double sc, double sb, ocmp Ua, ocmp Uc, ocmp *Ub, double nx, double ny, double nz) {
  const double Uan = Ua.x * nx + Ua.y * ny + Ua.z * nz;
  const double Ucn = Uc.x * nx + Uc.y * ny + Uc.z * nz;
  const double Ubn = Uan + (Ucn - Uan) * sb / sc;
  const double ut = Uc.x - Ucn * nx, vt = Uc.y - Ucn * ny, wt = Uc.z - Ucn * nz;
  Ub->x = ut; Ub->y = vt; Ub->z = wt;
  Ub->x += Ubn * nx;
  Ub->y += Ubn * ny;
  Ub->z += Ubn * nz;
}

void o_ibm_interpolation(octx *c) {
  if (c->ib_n <= 0) return;
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const ofield *nv = &L->nvert, *icsi = &L->csi, *jeta = &L->eta, *kzet = &L->zet;
  ofield *ucat = &c->ucat, *ucont = &c->ucont, *p = &c->p;
  ofield lucat, lp;
  of_alloc(&lucat, mx, my, mz, 3);
  of_alloc(&lp, mx, my, mz, 1);
  of_copy(&lp, p);                    /* lP: P at entry, refreshed only after the sweeps */
  const int itr_tmp = 6;
  for (int tmp = 0; tmp < itr_tmp; tmp++) {
    of_copy(&lucat, ucat);            /* lUcat: Ucat at the sweep's start */
    for (int q = 0; q < c->ib_n; q++) {
      const int *I = c->ib_idx + 12 * q;
      const double *W = c->ib_val + 12 * q;
      const int i = I[0], j = I[1], k = I[2];
      const double sk1 = W[0], sk2 = W[1], sk3 = W[2];
      const double sb = W[3], sc = sb + W[4];
      const ocmp Ua = {W[5], W[6], W[7]};
      const double dun = W[8];
      const ocmp *a1 = &V(&lucat, I[5], I[4], I[3]), *a2 = &V(&lucat, I[8], I[7], I[6]),
                 *a3 = &V(&lucat, I[11], I[10], I[9]);
      ocmp Uc;
      Uc.x = (a1->x * sk1 + a2->x * sk2 + a3->x * sk3);
      Uc.y = (a1->y * sk1 + a2->y * sk2 + a3->y * sk3);
      Uc.z = (a1->z * sk1 + a2->z * sk2 + a3->z * sk3);
      if (c->ib_mode == 1) ib_freeslip(sc, sb, Ua, Uc, &V(ucat, k, j, i), W[9], W[10], W[11]);
      else ib_noslip(sc, sb, Ua, Uc, &V(ucat, k, j, i));
      const double cv1 = S(&lp, I[5], I[4], I[3]), cv2 = S(&lp, I[8], I[7], I[6]), cv3 = S(&lp, I[11], I[10], I[9]);
      S(p, k, j, i) = (cv1 * sk1 + cv2 * sk2 + cv3 * sk3);
      S(p, k, j, i) += dun / c->dt * (sc - sb);
    }
    o_g2l(ucat, L->per);
    of_copy(&lucat, ucat);
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          ocmp uc;
          const ocmp *u0 = &V(&lucat, k, j, i), *ui = &V(&lucat, k, j, i + 1), *uj = &V(&lucat, k, j + 1, i),
                     *uk = &V(&lucat, k + 1, j, i);
          const ocmp *ci = &V(icsi, k, j, i), *cj = &V(jeta, k, j, i), *ck = &V(kzet, k, j, i);
          ocmp *f = &V(ucont, k, j, i);
          const int ib0 = (int)(S(nv, k, j, i) + 0.5) == 1;
          if (ib0 || (int)(S(nv, k, j, i + 1) + 0.5) == 1) {
            uc.x = (u0->x + ui->x) * 0.5; uc.y = (u0->y + ui->y) * 0.5; uc.z = (u0->z + ui->z) * 0.5;
            if (i < mx - 2) f->x = (uc.x * ci->x + uc.y * ci->y + uc.z * ci->z) * 1.0;
          }
          if (ib0 || (int)(S(nv, k, j + 1, i) + 0.5) == 1) {
            uc.x = (uj->x + u0->x) * 0.5; uc.y = (uj->y + u0->y) * 0.5; uc.z = (uj->z + u0->z) * 0.5;
            if (j < my - 2) f->y = (uc.x * cj->x + uc.y * cj->y + uc.z * cj->z) * 1.0;
          }
          if (ib0 || (int)(S(nv, k + 1, j, i) + 0.5) == 1) {
            uc.x = (uk->x + u0->x) * 0.5; uc.y = (uk->y + u0->y) * 0.5; uc.z = (uk->z + u0->z) * 0.5;
            if (k < mz - 2) f->z = (uc.x * ck->x + uc.y * ck->y + uc.z * ck->z) * 1.0;
          }
        }
    o_g2l(ucont, L->per);
    o_contra2cart(c);
  }
  o_g2l(p, L->per);
  of_free(&lucat); of_free(&lp);
}

/* ------------------------------------------------------------------ */
/* Convection (rhs.c:1463-1945), QUICK (central/LES branches omitted)   */
/* ------------------------------------------------------------------ */
static inline void quick(double um, double up, const ocmp *a2, const ocmp *a1, const ocmp *a0,
                         const ocmp *am, ocmp *fp) {
  /* fp = um*(coef*(-u[+2] - 2u[+1] + 3u[0]) + u[+1]) + up*(coef*(-u[-1] - 2u[0] + 3u[+1]) + u[0])
   * with a2=u[+2], a1=u[+1], a0=u[0], am=u[-1] */
  const double coef = 0.125;
  fp->x = um * (coef * (-a2->x - 2. * a1->x + 3. * a0->x) + a1->x) +
          up * (coef * (-am->x - 2. * a0->x + 3. * a1->x) + a0->x);
  fp->y = um * (coef * (-a2->y - 2. * a1->y + 3. * a0->y) + a1->y) +
          up * (coef * (-am->y - 2. * a0->y + 3. * a1->y) + a0->y);
  fp->z = um * (coef * (-a2->z - 2. * a1->z + 3. * a0->z) + a1->z) +
          up * (coef * (-am->z - 2. * a0->z + 3. * a1->z) + a0->z);
}

static void convection(octx *c, ofield *conv) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  ofield *u = &c->ucat, *nv = &L->nvert; const ofield *uc = &c->ucont;
  const double innerblank = 7.;
  const int sw = (L->per[0] || L->per[1] || L->per[2]) ? 3 : 2;
  const int gxs = L->per[0] ? -sw : 0, gxe = L->per[0] ? mx + sw : mx;
  const int gys = L->per[1] ? -sw : 0, gye = L->per[1] ? my + sw : my;
  const int gzs = L->per[2] ? -sw : 0, gze = L->per[2] ? mz + sw : mz;
  /* periodic far ghosts for QUICK (rhs.c:1557-1628) */
  if (L->bctype[0] == 7 || L->bctype[1] == 7) {
    for (int k = gzs; k < gze; k++) for (int j = gys; j < gye; j++) {
      V(u,k,j,-1) = V(u,k,j,-3); S(nv,k,j,-1) = S(nv,k,j,-3);
    }
    for (int k = gzs; k < gze; k++) for (int j = gys; j < gye; j++) {
      V(u,k,j,mx) = V(u,k,j,mx+2); S(nv,k,j,mx) = S(nv,k,j,mx+2);
    }
  }
  if (L->bctype[4] == 7 || L->bctype[5] == 7) {
    for (int j = gys; j < gye; j++) for (int i = gxs; i < gxe; i++) {
      V(u,-1,j,i) = V(u,-3,j,i); S(nv,-1,j,i) = S(nv,-3,j,i);
    }
    for (int j = gys; j < gye; j++) for (int i = gxs; i < gxe; i++) {
      V(u,mz,j,i) = V(u,mz+2,j,i); S(nv,mz,j,i) = S(nv,mz+2,j,i);
    }
  }
  ofield fp1, fp2, fp3;
  of_alloc(&fp1, mx, my, mz, 3); of_alloc(&fp2, mx, my, mz, 3); of_alloc(&fp3, mx, my, mz, 3);
#define FL(n) ((n) < 0.1 || (n) > innerblank)
  /* i direction (rhs.c:1636-1706) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 0; i < mx - 1; i++) {
        double ucon = V(uc,k,j,i).x * 0.5, up = ucon + fabs(ucon), um = ucon - fabs(ucon);
        ocmp *f = &V(&fp1,k,j,i);
        if (i > 0 && i < mx - 2 && FL(S(nv,k,j,i+1)) && FL(S(nv,k,j,i-1)))
          quick(um, up, &V(u,k,j,i+2), &V(u,k,j,i+1), &V(u,k,j,i), &V(u,k,j,i-1), f);
        else if (i == 0 || S(nv,k,j,i-1) > 0.1) {
          if (L->bctype[0] == 7 && i == 0 && (S(nv,k,j,i-1) < 0.1 && S(nv,k,j,i+1) < 0.1))
            quick(um, up, &V(u,k,j,i+2), &V(u,k,j,i+1), &V(u,k,j,i), &V(u,k,j,i-1), f);
          else
            quick(um, up, &V(u,k,j,i+2), &V(u,k,j,i+1), &V(u,k,j,i), &V(u,k,j,i), f);
        } else if (i == mx - 2 || S(nv,k,j,i+1) > 0.1) {
          if (L->bctype[0] == 7 && i == mx - 2 && (S(nv,k,j,i-1) < 0.1 && S(nv,k,j,i+1) < 0.1))
            quick(um, up, &V(u,k,j,i+2), &V(u,k,j,i+1), &V(u,k,j,i), &V(u,k,j,i-1), f);
          else
            quick(um, up, &V(u,k,j,i+1), &V(u,k,j,i+1), &V(u,k,j,i), &V(u,k,j,i-1), f);
        }
      }
  /* j direction (rhs.c:1709-1800) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 0; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        double ucon = V(uc,k,j,i).y * 0.5, up = ucon + fabs(ucon), um = ucon - fabs(ucon);
        ocmp *f = &V(&fp2,k,j,i);
        if (j > 0 && j < my - 2 && FL(S(nv,k,j+1,i)) && FL(S(nv,k,j-1,i)))
          quick(um, up, &V(u,k,j+2,i), &V(u,k,j+1,i), &V(u,k,j,i), &V(u,k,j-1,i), f);
        else if (j == 0 || S(nv,k,j-1,i) > 0.1)
          quick(um, up, &V(u,k,j+2,i), &V(u,k,j+1,i), &V(u,k,j,i), &V(u,k,j,i), f);
        else if (j == my - 2 || S(nv,k,j+1,i) > 0.1)
          quick(um, up, &V(u,k,j+1,i), &V(u,k,j+1,i), &V(u,k,j,i), &V(u,k,j-1,i), f);
      }
  /* k direction (rhs.c:1803-1897) */
  for (int k = 0; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        double ucon = V(uc,k,j,i).z * 0.5, up = ucon + fabs(ucon), um = ucon - fabs(ucon);
        ocmp *f = &V(&fp3,k,j,i);
        if (k > 0 && k < mz - 2 && FL(S(nv,k+1,j,i)) && FL(S(nv,k-1,j,i)))
          quick(um, up, &V(u,k+2,j,i), &V(u,k+1,j,i), &V(u,k,j,i), &V(u,k-1,j,i), f);
        else if (k < mz - 2 && (k == 0 || S(nv,k-1,j,i) > 0.1)) {
          if (L->bctype[4] == 7 && k == 0 && (S(nv,k-1,j,i) < 0.1 && S(nv,k+1,j,i) < 0.1))
            quick(um, up, &V(u,k+2,j,i), &V(u,k+1,j,i), &V(u,k,j,i), &V(u,k-1,j,i), f);
          else
            quick(um, up, &V(u,k+2,j,i), &V(u,k+1,j,i), &V(u,k,j,i), &V(u,k,j,i), f);
        } else if (k > 0 && (k == mz - 2 || S(nv,k+1,j,i) > 0.1)) {
          if (L->bctype[4] == 7 && k == mz - 2 && (S(nv,k-1,j,i) < 0.1 && S(nv,k+1,j,i) < 0.1))
            quick(um, up, &V(u,k+2,j,i), &V(u,k+1,j,i), &V(u,k,j,i), &V(u,k-1,j,i), f);
          else
            quick(um, up, &V(u,k+1,j,i), &V(u,k+1,j,i), &V(u,k,j,i), &V(u,k-1,j,i), f);
        }
      }
#undef FL
  of_zero(conv);
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ocmp *cv = &V(conv,k,j,i);
        cv->x = V(&fp1,k,j,i).x - V(&fp1,k,j,i-1).x + V(&fp2,k,j,i).x - V(&fp2,k,j-1,i).x +
                V(&fp3,k,j,i).x - V(&fp3,k-1,j,i).x;
        cv->y = V(&fp1,k,j,i).y - V(&fp1,k,j,i-1).y + V(&fp2,k,j,i).y - V(&fp2,k,j-1,i).y +
                V(&fp3,k,j,i).y - V(&fp3,k-1,j,i).y;
        cv->z = V(&fp1,k,j,i).z - V(&fp1,k,j,i-1).z + V(&fp2,k,j,i).z - V(&fp2,k,j-1,i).z +
                V(&fp3,k,j,i).z - V(&fp3,k-1,j,i).z;
      }
  of_free(&fp1); of_free(&fp2); of_free(&fp3);
}

/* ------------------------------------------------------------------ */
/* Viscous (rhs.c:1947-2565), laminar                                   */
/* ------------------------------------------------------------------ */
/* viscous flux through one face given the nine velocity gradients and the
 * face metrics (a = csi, e = eta, z = zet of the face), normal axis d:
 * fp = (g1*du/dc + g2*du/de + g3*du/dz + r.n) * ajc * nu  (rhs.c:2148-2165) */
static void visc_flux(int d, const double du[3][3], const ocmp *a, const ocmp *e, const ocmp *z,
                      double ajc, double nu, ocmp *fp) {
  double csi0 = a->x, csi1 = a->y, csi2 = a->z;
  double eta0 = e->x, eta1 = e->y, eta2 = e->z;
  double zet0 = z->x, zet1 = z->y, zet2 = z->z;
  double n0, n1, n2, g11, g21, g31;
  if (d == 0) { n0 = csi0; n1 = csi1; n2 = csi2;
    g11 = csi0 * csi0 + csi1 * csi1 + csi2 * csi2; g21 = eta0 * csi0 + eta1 * csi1 + eta2 * csi2;
    g31 = zet0 * csi0 + zet1 * csi1 + zet2 * csi2; }
  else if (d == 1) { n0 = eta0; n1 = eta1; n2 = eta2;
    g11 = csi0 * eta0 + csi1 * eta1 + csi2 * eta2; g21 = eta0 * eta0 + eta1 * eta1 + eta2 * eta2;
    g31 = zet0 * eta0 + zet1 * eta1 + zet2 * eta2; }
  else { n0 = zet0; n1 = zet1; n2 = zet2;
    g11 = csi0 * zet0 + csi1 * zet1 + csi2 * zet2; g21 = eta0 * zet0 + eta1 * zet1 + eta2 * zet2;
    g31 = zet0 * zet0 + zet1 * zet1 + zet2 * zet2; }
  /* du[comp][dir]: comp u,v,w; dir c,e,z */
  double dudc = du[0][0], dude = du[0][1], dudz = du[0][2];
  double dvdc = du[1][0], dvde = du[1][1], dvdz = du[1][2];
  double dwdc = du[2][0], dwde = du[2][1], dwdz = du[2][2];
  double r11 = dudc * csi0 + dude * eta0 + dudz * zet0;
  double r21 = dvdc * csi0 + dvde * eta0 + dvdz * zet0;
  double r31 = dwdc * csi0 + dwde * eta0 + dwdz * zet0;
  double r12 = dudc * csi1 + dude * eta1 + dudz * zet1;
  double r22 = dvdc * csi1 + dvde * eta1 + dvdz * zet1;
  double r32 = dwdc * csi1 + dwde * eta1 + dwdz * zet1;
  double r13 = dudc * csi2 + dude * eta2 + dudz * zet2;
  double r23 = dvdc * csi2 + dvde * eta2 + dvdz * zet2;
  double r33 = dwdc * csi2 + dwde * eta2 + dwdz * zet2;
  fp->x = 0; fp->y = 0; fp->z = 0;
  fp->x += (g11 * dudc + g21 * dude + g31 * dudz + r11 * n0 + r21 * n1 + r31 * n2) * ajc * (nu);
  fp->y += (g11 * dvdc + g21 * dvde + g31 * dvdz + r12 * n0 + r22 * n1 + r32 * n2) * ajc * (nu);
  fp->z += (g11 * dwdc + g21 * dwde + g31 * dwdz + r13 * n0 + r23 * n1 + r33 * n2) * ajc * (nu);
}

/* tangential velocity derivative on a face (rhs.c:2070-2135): one-sided
 * toward the other pair when a neighbour cell is an IB/solid node */
static void tangent_grad(const ofield *u, const ofield *nv, int d, int t, int i, int j, int k,
                         double out[3]) {
  const double solid = 0.5, innerblank = 7.;
  int o[3] = {0, 0, 0}; o[d] = 1;
  int up[3] = {0, 0, 0}, dn[3] = {0, 0, 0}; up[t] = 1; dn[t] = -1;
#define NV(di, dj, dk) S(nv, k + (dk), j + (dj), i + (di))
#define UU(di, dj, dk) V(u, k + (dk), j + (dj), i + (di))
#define SB(n) ((n) > solid && (n) < innerblank)
  /* pairs: the two cells of the face at tangential offset +1 and -1 */
  int hi = SB(NV(up[0], up[1], up[2])) || SB(NV(up[0] + o[0], up[1] + o[1], up[2] + o[2]));
  int lo = SB(NV(dn[0], dn[1], dn[2])) || SB(NV(dn[0] + o[0], dn[1] + o[1], dn[2] + o[2]));
  const ocmp *A, *B, *C, *D; double f;
  if (hi) {        /* (u[0]+u[o] - u[-1]-u[o-1]) * 0.5 */
    A = &UU(o[0], o[1], o[2]); B = &UU(0, 0, 0);
    C = &UU(o[0] + dn[0], o[1] + dn[1], o[2] + dn[2]); D = &UU(dn[0], dn[1], dn[2]); f = 0.5;
  } else if (lo) { /* (u[+1]+u[o+1] - u[0]-u[o]) * 0.5 */
    A = &UU(o[0] + up[0], o[1] + up[1], o[2] + up[2]); B = &UU(up[0], up[1], up[2]);
    C = &UU(o[0], o[1], o[2]); D = &UU(0, 0, 0); f = 0.5;
  } else {         /* (u[+1]+u[o+1] - u[-1]-u[o-1]) * 0.25 */
    A = &UU(o[0] + up[0], o[1] + up[1], o[2] + up[2]); B = &UU(up[0], up[1], up[2]);
    C = &UU(o[0] + dn[0], o[1] + dn[1], o[2] + dn[2]); D = &UU(dn[0], dn[1], dn[2]); f = 0.25;
  }
  out[0] = (A->x + B->x - C->x - D->x) * f;
  out[1] = (A->y + B->y - C->y - D->y) * f;
  out[2] = (A->z + B->z - C->z - D->z) * f;
#undef NV
#undef UU
#undef SB
}

static void viscous(octx *c, ofield *visc) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const ofield *u = &c->ucat, *nv = &L->nvert;
  const double nu = 1. / c->ren;
  ofield fp[3];
  for (int d = 0; d < 3; d++) of_alloc(&fp[d], mx, my, mz, 3);
  const ofield *fa[3][3] = {{&L->icsi, &L->ieta, &L->izet}, {&L->jcsi, &L->jeta, &L->jzet},
                            {&L->kcsi, &L->keta, &L->kzet}};
  const ofield *faj[3] = {&L->iaj, &L->jaj, &L->kaj};
  for (int d = 0; d < 3; d++) {
    int k0 = d == 2 ? 0 : 1, j0 = d == 1 ? 0 : 1, i0 = d == 0 ? 0 : 1;
    for (int k = k0; k < mz - 1; k++)
      for (int j = j0; j < my - 1; j++)
        for (int i = i0; i < mx - 1; i++) {
          double du[3][3];   /* du[comp][dir] */
          int o[3] = {0, 0, 0}; o[d] = 1;
          for (int t = 0; t < 3; t++) {
            double g[3];
            if (t == d) {
              const ocmp *a = &V(u, k + o[2], j + o[1], i + o[0]), *b = &V(u, k, j, i);
              g[0] = a->x - b->x; g[1] = a->y - b->y; g[2] = a->z - b->z;
            } else tangent_grad(u, nv, d, t, i, j, k, g);
            du[0][t] = g[0]; du[1][t] = g[1]; du[2][t] = g[2];
          }
          visc_flux(d, du, &V(fa[d][0], k, j, i), &V(fa[d][1], k, j, i), &V(fa[d][2], k, j, i),
                    S(faj[d], k, j, i), nu, &V(&fp[d], k, j, i));
        }
  }
  of_zero(visc);
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ocmp *v = &V(visc,k,j,i);
        v->x = (V(&fp[0],k,j,i).x - V(&fp[0],k,j,i-1).x + V(&fp[1],k,j,i).x - V(&fp[1],k,j-1,i).x +
                V(&fp[2],k,j,i).x - V(&fp[2],k-1,j,i).x);
        v->y = (V(&fp[0],k,j,i).y - V(&fp[0],k,j,i-1).y + V(&fp[1],k,j,i).y - V(&fp[1],k,j-1,i).y +
                V(&fp[2],k,j,i).y - V(&fp[2],k-1,j,i).y);
        v->z = (V(&fp[0],k,j,i).z - V(&fp[0],k,j,i-1).z + V(&fp[1],k,j,i).z - V(&fp[1],k,j-1,i).z +
                V(&fp[2],k,j,i).z - V(&fp[2],k-1,j,i).z);
      }
  for (int d = 0; d < 3; d++) of_free(&fp[d]);
}

/* ------------------------------------------------------------------ */
/* FormFunction1 (rhs.c:2567-3351)                                      */
/* ------------------------------------------------------------------ */
typedef struct { double dpdc, dpde, dpdz; long drops; } ff1_carry;

/* tangential pressure derivative with the FF1 semantics: DROP keeps *v */
static void ff1_tangent(const olevel *L, const ofield *p, int d, int t, int i, int j, int k,
                        double *v, long *drops) {
  const ofield *nv = &L->nvert;
  int o[3] = {0, 0, 0}; o[d] = 1;
  int up[3] = {0, 0, 0}, dn[3] = {0, 0, 0}; up[t] = 1; dn[t] = -1;
  int m_ax[3] = {L->mx, L->my, L->mz}, ijk[3] = {i, j, k};
  double hi = S(nv, k + up[2], j + up[1], i + up[0]) + S(nv, k + up[2] + o[2], j + up[1] + o[1], i + up[0] + o[0]);
  double lo = S(nv, k + dn[2], j + dn[1], i + dn[0]) + S(nv, k + dn[2] + o[2], j + dn[1] + o[1], i + dn[0] + o[0]);
  int per = (t == 0) ? (L->bctype[0] == 7) : (t == 1) ? (L->bctype[2] == 7) : (L->bctype[4] == 7);
  int mode = o_xmode(hi, lo, ijk[t] == m_ax[t] - 2, ijk[t] == 1, per);
#define PP(di, dj, dk) S(p, k + (dk), j + (dj), i + (di))
  if (mode == 1)
    *v = (PP(0,0,0) + PP(o[0],o[1],o[2]) - PP(dn[0],dn[1],dn[2]) - PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.5;
  else if (mode == 2)
    *v = (PP(up[0],up[1],up[2]) + PP(o[0]+up[0],o[1]+up[1],o[2]+up[2]) - PP(0,0,0) - PP(o[0],o[1],o[2])) * 0.5;
  else if (mode == 0)
    *v = (PP(up[0],up[1],up[2]) + PP(o[0]+up[0],o[1]+up[1],o[2]+up[2]) - PP(dn[0],dn[1],dn[2]) - PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.25;
  else {
    /* DROP: the variable keeps its previous value; count it when the face
     * is not masked afterwards (both cells fluid) */
    if (S(nv, k, j, i) <= 0.1 && S(nv, k + o[2], j + o[1], i + o[0]) <= 0.1) (*drops)++;
  }
#undef PP
}

static void ff1_face(octx *c, int d, int i, int j, int k, const ofield *rct, ofield *rhs, ff1_carry *cr) {
  olevel *L = &c->lev[c->nlev - 1];
  const ofield *p = &c->p;
  int o[3] = {0, 0, 0}; o[d] = 1;
  double *dv[3] = {&cr->dpdc, &cr->dpde, &cr->dpdz};
  for (int t = 0; t < 3; t++) {
    if (t == d) *dv[t] = S(p, k + o[2], j + o[1], i + o[0]) - S(p, k, j, i);
    else ff1_tangent(L, p, d, t, i, j, k, dv[t], &cr->drops);
  }
  const ocmp *cs, *et, *ze; double aj;
  const ocmp *nrm;
  if (d == 0) { cs = &V(&L->icsi,k,j,i); et = &V(&L->ieta,k,j,i); ze = &V(&L->izet,k,j,i); aj = S(&L->iaj,k,j,i); nrm = cs; }
  else if (d == 1) { cs = &V(&L->jcsi,k,j,i); et = &V(&L->jeta,k,j,i); ze = &V(&L->jzet,k,j,i); aj = S(&L->jaj,k,j,i); nrm = et; }
  else { cs = &V(&L->kcsi,k,j,i); et = &V(&L->keta,k,j,i); ze = &V(&L->kzet,k,j,i); aj = S(&L->kaj,k,j,i); nrm = ze; }
  const double *rc0 = &V(rct, k, j, i).x, *rc1 = &V(rct, k + o[2], j + o[1], i + o[0]).x;
  double r = 0.5 * (rc0[d] + rc1[d]);
  r -= (cr->dpdc * (cs->x * nrm->x + cs->y * nrm->y + cs->z * nrm->z) +
        cr->dpde * (et->x * nrm->x + et->y * nrm->y + et->z * nrm->z) +
        cr->dpdz * (ze->x * nrm->x + ze->y * nrm->y + ze->z * nrm->z)) * aj;
  (&V(rhs, k, j, i).x)[d] = r;
}

void o_formfunction1(octx *c, ofield *rhs) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  ofield conv, visc, rct;
  of_alloc(&conv, mx, my, mz, 3); of_alloc(&visc, mx, my, mz, 3); of_alloc(&rct, mx, my, mz, 3);
  o_contra2cart(c);
  convection(c, &conv);
  viscous(c, &visc);
  const ofield *csi = &L->csi, *eta = &L->eta, *zet = &L->zet, *aj = &L->aj;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ocmp rc;   /* Rc = Visc - Conv (VecWAXPY, rhs.c:2663) */
        rc.x = -1.0 * V(&conv,k,j,i).x + V(&visc,k,j,i).x;
        rc.y = -1.0 * V(&conv,k,j,i).y + V(&visc,k,j,i).y;
        rc.z = -1.0 * V(&conv,k,j,i).z + V(&visc,k,j,i).z;
        double a = S(aj, k, j, i);
        V(&rct,k,j,i).x = a * (0.5 * (V(csi,k,j,i).x + V(csi,k,j,i-1).x) * rc.x +
                               0.5 * (V(csi,k,j,i).y + V(csi,k,j,i-1).y) * rc.y +
                               0.5 * (V(csi,k,j,i).z + V(csi,k,j,i-1).z) * rc.z);
        V(&rct,k,j,i).y = a * (0.5 * (V(eta,k,j,i).x + V(eta,k,j-1,i).x) * rc.x +
                               0.5 * (V(eta,k,j,i).y + V(eta,k,j-1,i).y) * rc.y +
                               0.5 * (V(eta,k,j,i).z + V(eta,k,j-1,i).z) * rc.z);
        V(&rct,k,j,i).z = a * (0.5 * (V(zet,k,j,i).x + V(zet,k-1,j,i).x) * rc.x +
                               0.5 * (V(zet,k,j,i).y + V(zet,k-1,j,i).y) * rc.y +
                               0.5 * (V(zet,k,j,i).z + V(zet,k-1,j,i).z) * rc.z);
      }
  o_g2l(&rct, L->per);      /* DMLocalToLocal (rhs.c:2720) */
  if (L->bctype[0] == 7)
    for (int k = 1; k < mz - 1; k++) for (int j = 1; j < my - 1; j++) {
      V(&rct,k,j,0).x = V(&rct,k,j,-2).x; V(&rct,k,j,mx-1).x = V(&rct,k,j,mx+1).x;
    }
  if (L->bctype[4] == 7)
    for (int j = 1; j < my - 1; j++) for (int i = 1; i < mx - 1; i++) {
      V(&rct,0,j,i).z = V(&rct,-2,j,i).z; V(&rct,mz-1,j,i).z = V(&rct,mz+1,j,i).z;
    }
  ff1_carry cr = {0., 0., 0., 0};
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ff1_face(c, 0, i, j, k, &rct, rhs, &cr);
        ff1_face(c, 1, i, j, k, &rct, rhs, &cr);
        ff1_face(c, 2, i, j, k, &rct, rhs, &cr);
      }
  if (L->bctype[0] == 7)
    for (int k = 1; k < mz - 1; k++) for (int j = 1; j < my - 1; j++) ff1_face(c, 0, 0, j, k, &rct, rhs, &cr);
  if (L->bctype[4] == 7)
    for (int j = 1; j < my - 1; j++) for (int i = 1; i < mx - 1; i++) ff1_face(c, 2, i, j, 0, &rct, rhs, &cr);
  c->drop_count += cr.drops;
  /* masks (rhs.c:3280-3306) */
  const ofield *nv = &L->nvert;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ocmp *r = &V(rhs,k,j,i);
        if (S(nv,k,j,i) > 0.1) { r->x = 0; r->y = 0; r->z = 0; }
        if (S(nv,k,j,i+1) > 0.1) r->x = 0;
        if (S(nv,k,j+1,i) > 0.1) r->y = 0;
        if (S(nv,k+1,j,i) > 0.1) r->z = 0;
      }
  of_free(&conv); of_free(&visc); of_free(&rct);
}

/* ------------------------------------------------------------------ */
/* RungeKutta (implicitsolver.c:32-119)                                 */
/* ------------------------------------------------------------------ */
/* InflowFlux, OutflowFlux, FormBCS: the boundary update the reference
 * runs before and after every RK stage (implicitsolver.c:43-46, 79-84,
 * 7125-7127, 7161-7164) */
static void bc_update(octx *c) {
  o_inflowflux(c);
  o_outflowflux(c);
  o_formbcs(c);
}

void o_rungekutta(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  bc_update(c);
  o_ibm_interpolation(c);           /* implicitsolver.c:47-51 */
  of_zero(&c->rhs);
  const int mx = L->mx, my = L->my, mz = L->mz;
  for (int istage = 0; istage < 4; istage++) {
    o_formfunction1(c, &c->rhs);
    double a = alfa[istage] * c->dt * c->st;
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          ocmp *u = &V(&c->ucont,k,j,i); const ocmp *uo = &V(&c->ucont_o,k,j,i), *r = &V(&c->rhs,k,j,i);
          u->x = a * r->x + uo->x; u->y = a * r->y + uo->y; u->z = a * r->z + uo->z;
        }
    wall_flux_zero(c);
    o_g2l(&c->ucont, L->per);
    bc_update(c);
  }
  o_ibm_interpolation(c);           /* after the stages, implicitsolver.c:87-90 */
}

/* ------------------------------------------------------------------ */
/* CalcRHS (implicitsolver.c:6449-6741): FormFunction1 minus the BDF    */
/* time derivative, then the explicit boundary masks                     */
/* ------------------------------------------------------------------ */
void o_calcrhs(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const int mx = L->mx, my = L->my, mz = L->mz;
  const int xper = L->bctype[0] == 7, yper = L->bctype[2] == 7, zper = L->bctype[4] == 7;
  ofield *r = &c->rhs, *nvf = &L->nvert;
  o_formfunction1(c, r);
  /* du = 1.5 u - 2 u_o + 0.5 u_rm1 except on the first step of a run
   * (COEF_TIME_ACCURACY = 1.5), then Rhs += (-1/dt) du (6497-6511) */
  const int bdf2 = c->steps_done > 0;
  const double idt = -1. / c->dt;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const ocmp *u = &V(&c->ucont, k, j, i), *uo = &V(&c->ucont_o, k, j, i), *um = &V(&c->ucont_rm1, k, j, i);
        ocmp *rr = &V(r, k, j, i);
        double d[3];
        const double uu[3] = {u->x, u->y, u->z}, oo[3] = {uo->x, uo->y, uo->z}, mm[3] = {um->x, um->y, um->z};
        for (int q = 0; q < 3; q++) {
          if (bdf2) { d[q] = uu[q] * COEF_TIME_ACCURACY; d[q] = d[q] + (-2.) * oo[q]; d[q] = d[q] + 0.5 * mm[q]; }
          else d[q] = -1. * oo[q] + uu[q];
        }
        rr->x = rr->x + idt * d[0]; rr->y = rr->y + idt * d[1]; rr->z = rr->z + idt * d[2];
      }
#define NV(kk, jj, ii) S(nvf, kk, jj, ii)
  /* interior faces (6524-6538): IB neighbours and, on a non-periodic axis,
   * the last face (i > mx-3 etc.; mz_end = mz-3 without the outlet type 8) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        ocmp *rr = &V(r, k, j, i);
        if (NV(k, j, i) + NV(k, j, i + 1) > 0.1 || (i > mx - 3 && !xper)) rr->x = 0.;
        if (NV(k, j, i) + NV(k, j + 1, i) > 0.1 || (j > my - 3 && !yper)) rr->y = 0.;
        if (NV(k, j, i) + NV(k + 1, j, i) > 0.1 || (k > mz - 3 && !zper)) rr->z = 0.;
      }
  /* boundary planes (6541-6690); every branch only zeroes entries, so the
   * passes commute.  In the periodic upper-plane branches the reference
   * indexes with a stale loop variable: the first pass repeats the lower
   * plane's masks, the rest act on the plane that is zeroed right after. */
  for (int k = 0; k < mz - 1; k++)
    for (int j = 0; j < my - 1; j++) {
      ocmp *a = &V(r, k, j, 0);
      if (!xper) { a->x = a->y = a->z = 0.; }
      else {
        if (NV(k, j, 0) + NV(k, j, 1) > 0.1) a->x = 0.;
        if (NV(k, j, 0) + NV(k, j + 1, 0) > 0.1) a->y = 0.;
        if (NV(k, j, 0) + NV(k + 1, j, 0) > 0.1) a->z = 0.;
      }
      if (!xper) V(r, k, j, mx - 2).x = 0.;
      ocmp *b = &V(r, k, j, mx - 1);
      b->x = b->y = b->z = 0.;
    }
  for (int k = 0; k < mz - 1; k++)
    for (int i = 0; i < mx - 1; i++) {
      ocmp *a = &V(r, k, 0, i);
      if (!yper) { a->x = a->y = a->z = 0.; }
      else {
        if (NV(k, 0, i) + NV(k, 0, i + 1) > 0.1) a->x = 0.;
        if (NV(k, 0, i) + NV(k, 1, i) > 0.1) a->y = 0.;
        if (NV(k, 0, i) + NV(k + 1, 0, i) > 0.1) a->z = 0.;
      }
      if (!yper) V(r, k, my - 2, i).y = 0.;
      ocmp *b = &V(r, k, my - 1, i);
      b->x = b->y = b->z = 0.;
    }
  for (int j = 0; j < my - 1; j++)
    for (int i = 0; i < mx - 1; i++) {
      ocmp *a = &V(r, 0, j, i);
      if (!zper) { a->x = a->y = a->z = 0.; }
      else {
        if (NV(0, j, i) + NV(0, j, i + 1) > 0.1) a->x = 0.;
        if (NV(0, j, i) + NV(0, j + 1, i) > 0.1) a->y = 0.;
        if (NV(0, j, i) + NV(1, j, i) > 0.1) a->z = 0.;
      }
      if (!zper) V(r, mz - 2, j, i).z = 0.;
      ocmp *b = &V(r, mz - 1, j, i);
      b->x = b->y = b->z = 0.;
    }
#undef NV
  o_g2l(r, L->per);
}

/* max |a - b| over the storage points of two Cmpnts fields (b NULL: |a|) */
static double absmax3(const octx *c, const ofield *a, const ofield *b) {
  const olevel *L = &c->lev[c->nlev - 1];
  double m = 0.;
  for (int k = 0; k < L->mz; k++)
    for (int j = 0; j < L->my; j++)
      for (int i = 0; i < L->mx; i++) {
        const ocmp *p = &V(a, k, j, i);
        double v[3] = {p->x, p->y, p->z};
        if (b) { const ocmp *q = &V(b, k, j, i); v[0] = q->x - p->x; v[1] = q->y - p->y; v[2] = q->z - p->z; }
        for (int t = 0; t < 3; t++) if (fabs(v[t]) > m) m = fabs(v[t]);
      }
  return m;
}

/* ImpRK (implicitsolver.c:7094-7256): dual time stepping, four pseudo-time
 * RK stages per pseudo iteration on CalcRHS, stopped on the infinity norm
 * of the pseudo-time change; lambda (initially -cfl) halves and the
 * iteration is rejected when both the residual and the change grow */
void o_imprk(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  bc_update(c);
  o_ibm_interpolation(c);           /* implicitsolver.c:7128-7131 */
  of_zero(&c->rhs);
  of_copy(&c->pucont, &c->ucont);
  o_calcrhs(c);
  double normF0 = absmax3(c, &c->rhs, NULL), normF1 = 1000., normdU1 = 1000.;
  double lambda = c->cfl;
  double normdU = 10., reldU = 1., normdU0 = 0., normdU_bk = 0., reldU_bk = 1., normF_bk = 0.;
  (void)normF0;
  int pseudot = 0;
  while (((normdU > c->imp_atol && reldU > c->imp_rtol) || pseudot < 1) && pseudot < c->imp_max_it) {
    pseudot++;
    for (int s = 0; s < 4; s++) {
      const double a = lambda * alfa[s] * c->dt * c->st;
      for (int k = 0; k < L->mz; k++)
        for (int j = 0; j < L->my; j++)
          for (int i = 0; i < L->mx; i++) {
            ocmp *u = &V(&c->ucont, k, j, i);
            const ocmp *pu = &V(&c->pucont, k, j, i), *r = &V(&c->rhs, k, j, i);
            u->x = a * r->x + pu->x; u->y = a * r->y + pu->y; u->z = a * r->z + pu->z;
          }
      o_g2l(&c->ucont, L->per);
      bc_update(c);
      o_calcrhs(c);
    }
    o_ibm_interpolation(c);         /* after the stages, implicitsolver.c:7170-7173 */
    normdU1 = normdU_bk;
    normdU_bk = absmax3(c, &c->ucont, &c->pucont);   /* dU = pU - U */
    if (pseudot == 1) { normdU0 = normdU_bk; reldU_bk = 1.; }
    else if (normdU0 > 1.e-10) reldU_bk = normdU_bk / normdU0;
    if (pseudot > 1) normF1 = normF_bk;
    normF_bk = absmax3(c, &c->rhs, NULL);
    if (pseudot == 1) normF0 = normF_bk;
    normdU = normdU_bk;
    reldU = reldU_bk;
    if (pseudot > 1 && normF_bk > normF1 && normdU_bk > normdU1 && lambda > 0.001) {
      lambda = 0.5 * lambda;
      pseudot--;
      of_copy(&c->ucont, &c->pucont);
      o_g2l(&c->ucont, L->per);
    } else {
      of_copy(&c->pucont, &c->ucont);
      normF1 = normF_bk;
      normdU1 = normdU_bk;
    }
  }
  c->imp_its = pseudot; c->imp_lambda = lambda; c->imp_normdU = normdU; c->imp_reldU = reldU;
}

/* ------------------------------------------------------------------ */
/* one fractional step (solvers.c:944-1156)                             */
/* ------------------------------------------------------------------ */
void o_flow_step(octx *c) {
  olevel *L = &c->lev[c->nlev - 1];
  /* U_rm1 <- U_o <- U (main.c:1131-1132), done here at the step start */
  if (c->steps_done > 0) of_copy(&c->ucont_rm1, &c->ucont_o);
  of_copy(&c->ucont_o, &c->ucont);          /* main.c:1127 / 1065 */
  double t0 = now_s();
  if (c->mom_scheme == 1) o_imprk(c);
  else o_rungekutta(c);
  double t1 = now_s();
  o_poisson_solve(c);
  double t2 = now_s();
  o_update_pressure(c);
  o_projection(c);
  o_g2l(&c->p, L->per);
  double t3 = now_s();
  o_divergence(c);
  o_ibm_interpolation(c);           /* solvers.c:1095-1101 */
  c->t_mom = t1 - t0; c->t_p = t2 - t1; c->t_pr = t3 - t2;
  c->steps_done++;
}
