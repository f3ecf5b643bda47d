/*
 * cv_oracle_sanitize.c -- TEST INFRASTRUCTURE: a stand-alone driver that runs
 * the CPU oracle's fractional steps under AddressSanitizer and
 * UndefinedBehaviorSanitizer (built by `make sanitize`, host code only).
 *
 * The reference path carries undefined behaviour of its own (a pressure
 * derivative carried over from an earlier face in FormFunction1,
 * rhs.c:2803-2861; a loop bound taken from a stale variable in ImpRK,
 * implicitsolver.c:7101,7198).  The oracle restates both with defined values;
 * this driver shows that the restatement itself reads and writes only what it
 * owns and computes nothing undefined over whole steps on immersed-boundary
 * cases, for the RK and ImpRK schemes.
 *
 * Usage: cv_oracle_sanitize IN OUT
 *   IN, little-endian: int32 hdr[16] = {IM, JM, KM, bctype[6], nlev, inlet,
 *     scheme, nsteps, n_ib, ib_mode, coarsen_auto}; double dhdr[6] = {Re, dt,
 *     x_c, y_c, z_c, rtol}; node coordinates IM*JM*KM*3; nvert mx*my*mz;
 *     Ucont mx*my*mz*3; IB indices n_ib*12 int32; IB values n_ib*12 double.
 *   OUT: p (mx*my*mz), Ucont (mx*my*mz*3), then per step {ksp its, maxdiv}.
 * Exit status 0 on success; the sanitizers abort on their first report.
 */
#include "cv_oracle.h"
#include <stdio.h>
#include <stdlib.h>

int o_put(octx *c, const char *name, int level, const double *in);
int o_get(octx *c, const char *name, int level, double *out);
void o_set_solver(octx *c, int restart, int max_it, double rtol, double atol, int smooth_it);
void o_stats(octx *c, double *out);

static void *slurp(FILE *f, size_t bytes, const char *what) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "cv_oracle_sanitize: short read of %s\n", what);
    exit(2);
  }
  return p;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  FILE *in = fopen(argv[1], "rb");
  if (!in) {
    perror(argv[1]);
    return 2;
  }
  int hdr[16];
  double dh[6];
  if (fread(hdr, sizeof hdr, 1, in) != 1 || fread(dh, sizeof dh, 1, in) != 1) {
    fprintf(stderr, "cv_oracle_sanitize: short header\n");
    return 2;
  }
  const int IM = hdr[0], JM = hdr[1], KM = hdr[2], nlev = hdr[9], inlet = hdr[10], scheme = hdr[11];
  const int nsteps = hdr[12], nib = hdr[13], ibmode = hdr[14], cauto = hdr[15];
  const size_t nn = (size_t)IM * JM * KM, ns = (size_t)(IM + 1) * (JM + 1) * (KM + 1);
  double *xyz = slurp(in, nn * 3 * sizeof(double), "coordinates");
  double *nv = slurp(in, ns * sizeof(double), "nvert");
  double *uc = slurp(in, ns * 3 * sizeof(double), "Ucont");
  int *ibi = slurp(in, (size_t)nib * 12 * sizeof(int), "IB indices");
  double *ibv = slurp(in, (size_t)nib * 12 * sizeof(double), "IB values");
  fclose(in);

  octx *c = o_create(IM, JM, KM, hdr + 3, nlev, 0, 0, 0, dh[0], dh[1], cauto);
  if (!c) {
    fprintf(stderr, "cv_oracle_sanitize: unsupported configuration\n");
    return 2;
  }
  o_set_inlet(c, inlet, dh[2], dh[3], dh[4]);
  o_set_coords(c, xyz);
  o_form_metrics_all(c);
  o_set_nvert(c, nv);
  o_nvert_restrict_all(c);
  if (nib) o_set_ib_nodes(c, nib, ibi, ibv, ibmode);
  o_put(c, "ucont", -1, uc);
  o_set_solver(c, 30, 200, dh[5], 0.0, 2);
  o_set_momentum(c, scheme, 20, 1e-7, 1e-4, 0.2);

  double *stat = calloc((size_t)nsteps * 2 + 1, sizeof(double));
  for (int s = 0; s < nsteps; s++) {
    o_flow_step(c);
    double st[16];
    o_stats(c, st);
    stat[2 * s] = st[0];
    stat[2 * s + 1] = st[4];
  }
  double *p = malloc(ns * sizeof(double));
  o_get(c, "p", -1, p);
  o_get(c, "ucont", -1, uc);
  FILE *out = fopen(argv[2], "wb");
  if (!out) {
    perror(argv[2]);
    return 2;
  }
  fwrite(p, sizeof(double), ns, out);
  fwrite(uc, sizeof(double), ns * 3, out);
  fwrite(stat, sizeof(double), (size_t)nsteps * 2, out);
  fclose(out);
  o_destroy(c);
  free(p);
  free(stat);
  free(xyz);
  free(nv);
  free(uc);
  free(ibi);
  free(ibv);
  return 0;
}
