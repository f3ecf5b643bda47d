/*
 * cv_oracle_api.c -- TEST INFRASTRUCTURE: flat accessors so that tests can
 * drive the CPU oracle through ctypes.  Nothing here is part of the product.
 */
#include "cv_oracle_int.h"
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static ofield *lookup(octx *c, const char *name, int level) {
  if (level < 0) level = c->nlev - 1;
  if (level >= c->nlev) return NULL;
  olevel *L = &c->lev[level];
#define F(nm, ptr) if (!strcmp(name, nm)) return ptr;
  F("ucont", &c->ucont) F("ucont_o", &c->ucont_o) F("ucont_rm1", &c->ucont_rm1) F("ucat", &c->ucat) F("p", &c->p)
  F("phi", &c->phi) F("rhs", &c->rhs) F("b", &c->b) F("ubcs", &c->ubcs)
  F("coor", &L->coor) F("csi", &L->csi) F("eta", &L->eta) F("zet", &L->zet) F("aj", &L->aj)
  F("icsi", &L->icsi) F("ieta", &L->ieta) F("izet", &L->izet) F("iaj", &L->iaj)
  F("jcsi", &L->jcsi) F("jeta", &L->jeta) F("jzet", &L->jzet) F("jaj", &L->jaj)
  F("kcsi", &L->kcsi) F("keta", &L->keta) F("kzet", &L->kzet) F("kaj", &L->kaj)
  F("cent", &L->cent) F("centx", &L->centx) F("centy", &L->centy) F("centz", &L->centz)
  F("gs", &L->gs) F("nvert", &L->nvert)
#undef F
  return NULL;
}

int o_get(octx *c, const char *name, int level, double *out) {
  ofield *f = lookup(c, name, level);
  if (!f) return -1;
  o_get_field(f, out);
  return 0;
}

int o_put(octx *c, const char *name, int level, const double *in) {
  ofield *f = lookup(c, name, level);
  if (!f) return -1;
  o_put_field(f, in);
  if (f == &c->ucont) c->steps_done = 0;          /* loaded fluxes start a run: BDF1 first */
  olevel *L = &c->lev[level < 0 ? c->nlev - 1 : level];
  o_g2l(f, L->per);
  return 0;
}

/* padded read (ghosts included): out has (m+2*OG)^3*dof entries */
int o_get_padded(octx *c, const char *name, int level, double *out) {
  ofield *f = lookup(c, name, level);
  if (!f) return -1;
  memcpy(out, f->a, (size_t)f->p[0] * f->p[1] * f->p[2] * f->dof * sizeof(double));
  return 0;
}

int o_dims(octx *c, int level, int *out6) {
  if (level < 0) level = c->nlev - 1;
  olevel *L = &c->lev[level];
  out6[0] = L->mx; out6[1] = L->my; out6[2] = L->mz;
  out6[3] = L->IM; out6[4] = L->JM; out6[5] = L->KM;
  return 0;
}

void o_set_solver(octx *c, int restart, int max_it, double rtol, double atol, int smooth_it) {
  c->restart = restart; c->max_it = max_it; c->rtol = rtol; c->atol = atol; c->smooth_it = smooth_it;
}

/* CSR export of a level's assembled operator (after o_poisson_lhs) */
int o_csr_nnz(octx *c, int level) {
  olevel *L = &c->lev[level];
  return L->rowptr ? L->rowptr[L->nrows] : 0;
}
int o_csr_get(octx *c, int level, int *rowptr, int *col, double *val) {
  olevel *L = &c->lev[level];
  if (!L->rowptr) return -1;
  memcpy(rowptr, L->rowptr, (L->nrows + 1) * sizeof(int));
  memcpy(col, L->col, L->rowptr[L->nrows] * sizeof(int));
  memcpy(val, L->val, L->rowptr[L->nrows] * sizeof(double));
  return 0;
}

void o_set_momentum(octx *c, int scheme, int imp_max_it, double imp_atol, double imp_rtol, double cfl) {
  c->mom_scheme = scheme; c->imp_max_it = imp_max_it; c->imp_atol = imp_atol; c->imp_rtol = imp_rtol; c->cfl = cfl;
}

void o_set_steps_done(octx *c, int n) { c->steps_done = n; }

void o_flux_stats(octx *c, double *out) {
  out[0] = c->flux_in_sum; out[1] = c->flux_out_sum; out[2] = c->flux_out_diag;
  out[3] = c->far_flux_in; out[4] = c->far_flux_out;
}

void o_imp_stats(octx *c, double *out) {
  out[0] = c->imp_its; out[1] = c->imp_lambda; out[2] = c->imp_normdU; out[3] = c->imp_reldU; out[4] = c->steps_done;
}

void o_stats(octx *c, double *out) {
  out[0] = c->ksp_its; out[1] = c->ksp_rnorm; out[2] = c->ksp_bnorm; out[3] = c->rhs_sum;
  out[4] = c->maxdiv; out[5] = (double)c->maxdiv_idx; out[6] = c->ibm_flux0; out[7] = c->ibm_area0;
  out[8] = c->ibm_flux1; out[9] = c->ibm_area1; out[10] = (double)c->drop_count;
  out[11] = c->t_mom; out[12] = c->t_p; out[13] = c->t_pr; out[14] = c->nthreads;
  out[15] = c->lev[0].multinullspace;
}

/* single-step helpers used by tests */
void o_rhs_into_b(octx *c) { o_poisson_rhs(c, &c->b); }
void o_ff1_into_rhs(octx *c) { o_formfunction1(c, &c->rhs); }
void o_lhs_all(octx *c) { for (int l = 0; l < c->nlev; l++) o_poisson_lhs(c, l); }
void o_matvec_level(octx *c, int level, const double *x, double *y) { o_csr_matvec(&c->lev[level], x, y); }

/* OpenMP threads of the following parallel regions (the block-Jacobi
 * smoother takes one block per thread, as PCBJACOBI takes one per rank) */
void o_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int o_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
