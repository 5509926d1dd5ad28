/*
 * orc_omp.h -- the oracle's OpenMP build (test infrastructure: the all-core
 * CPU baseline of the benchmark; liboracle_omp.so, -fopenmp -DORC_OMP).
 *
 * Per-edge and per-cell loops run in parallel.  The loops that scatter edge
 * terms into cells in edge order (CFL swept sums, the balance, the MUSCL
 * limiter) become per-cell gathers over the cell's incidences in ascending
 * edge order: each cell sees the same additions in the same order, so the
 * forward is the serial build's bit for bit.  The serial build (ORC_OMP
 * undefined) compiles none of this.
 */
#pragma once
#include <stdint.h>
#include <stdlib.h>

#ifdef ORC_OMP
#include <omp.h>
#define ORC_PAR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PAR
#endif

/* incidences of each cell, ascending: internal edges as (e << 1) | side
 * (side 1 = the cell is R), boundary edges as their index */
typedef struct {
    int64_t *iptr, *iinc, *bptr, *binc;
} orc_inc;

static inline void orc_inc_build(orc_inc *x, int64_t nc, int64_t ni, int64_t nb, const int64_t *L,
                                 const int64_t *R, const int64_t *bcell) {
    x->iptr = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    x->bptr = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    x->iinc = (int64_t *)calloc((size_t)(2 * ni + 1), sizeof(int64_t));
    x->binc = (int64_t *)calloc((size_t)(nb + 1), sizeof(int64_t));
    for (int64_t e = 0; e < ni; e++) {
        x->iptr[L[e] + 1]++;
        x->iptr[R[e] + 1]++;
    }
    for (int64_t e = 0; e < nb; e++) x->bptr[bcell[e] + 1]++;
    for (int64_t i = 0; i < nc; i++) {
        x->iptr[i + 1] += x->iptr[i];
        x->bptr[i + 1] += x->bptr[i];
    }
    int64_t *f = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    for (int64_t e = 0; e < ni; e++) {
        x->iinc[x->iptr[L[e]] + f[L[e]]++] = e << 1;
        x->iinc[x->iptr[R[e]] + f[R[e]]++] = (e << 1) | 1;
    }
    for (int64_t i = 0; i < nc; i++) f[i] = 0;
    for (int64_t e = 0; e < nb; e++) x->binc[x->bptr[bcell[e]] + f[bcell[e]]++] = e;
    free(f);
}

static inline void orc_inc_free(orc_inc *x) {
    free(x->iptr);
    free(x->iinc);
    free(x->bptr);
    free(x->binc);
}
