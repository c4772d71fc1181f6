/* Test infrastructure: index helpers shared by the oracle's C files. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include "biocfd_oracle.h"

#define ORC_PI 3.141592653589793115997963468544185161590576171875

/* Fortran (1-based, column-major) views of a block's arrays */
#define NX2(b) ((b)->nx + 2)
#define NY2(b) ((b)->ny + 2)
#define NZ2(b) ((b)->nz + 2)
#define IX(b, i, j, k) ((int64_t)(i) - 1 + NX2(b) * ((int64_t)(j) - 1 + NY2(b) * ((int64_t)(k) - 1)))
#define IXT(b, i, j, k) ((int64_t)(i) - 1 + ((b)->nx + 3) * ((int64_t)(j) - 1 + ((b)->ny + 3) * ((int64_t)(k) - 1)))
#define FLD(b, a, i, j, k) ((b)->a[IX(b, i, j, k)])
#define TAG3(b, a, i, j, k) ((b)->a[IXT(b, i, j, k)])
#define V1(a, i) ((a)[(int64_t)(i) - 1])
/* (n,3) coefficient arrays */
#define AC(a, n, i, c) ((a)[(int64_t)(i) - 1 + (int64_t)(n) * ((c) - 1)])
/* (count,3) index lists */
#define LIST(a, cnt, n, c) ((a)[(int64_t)(n) - 1 + (int64_t)(cnt) * ((c) - 1)])
/* (6, n+2) upwind tables */
#define TAB(a, r, i) ((a)[(r) - 1 + 6 * ((int64_t)(i) - 1)])

static inline double sq(double x) { return x * x; }

void *orc_xcalloc(size_t n, size_t sz);

#endif
