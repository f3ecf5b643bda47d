/* cv_oracle_int.h -- TEST INFRASTRUCTURE: private helpers of the CPU oracle. */
#ifndef CV_ORACLE_INT_H
#define CV_ORACLE_INT_H
#include "cv_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline size_t oix(const ofield *f, int i, int j, int k) {
  return ((size_t)(k + OG) * f->p[1] + (size_t)(j + OG)) * f->p[0] + (size_t)(i + OG);
}
#define S(f, k, j, i) ((f)->a[oix((f), (i), (j), (k))])
#define V(f, k, j, i) (((ocmp *)(f)->a)[oix((f), (i), (j), (k))])

void of_alloc(ofield *f, int mx, int my, int mz, int dof);
void of_free(ofield *f);
void of_zero(ofield *f);
void of_copy(ofield *dst, const ofield *src);

#define COEF_TIME_ACCURACY 1.5   /* variables.h:421 */

#endif
