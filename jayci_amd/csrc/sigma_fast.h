// Matrix-free sigma from per-string excitation tables: shared declarations.
#pragma once
#include "jci_internal.h"

// device view of one string set's excitation tables (see sigma_fast.hip)
struct ExcView {
    const long long *off1, *off2;
    const int2 *e1, *e2;
    const double *v1, *v2;
};

// kernel arguments of one space
struct FastArgs {
    const SpaceDesc *d;          // device descriptor
    const uint64_t *ma, *mb;     // string masks
    ExcView ta, tb;              // alpha / beta tables (the same when the sets coincide)
    const double *m2sq;          // square (P|Q), m1 x m1
    const double *hd;            // diagonal
    const long long *tile_off;   // per block first tile, nblk+1 (device)
    int m1;                      // orbital pairs n(n+1)/2
    int W;                       // beta strings per tile (64, 128 or 256)
    long long ntiles;
};

int jk_exc_count(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                 const unsigned long long *binom, long long *cnt1, long long *cnt2,
                 hipStream_t st);
int jk_exc_fill(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                const unsigned long long *binom, const double *m1, const double *m2,
                const long long *off1, const long long *off2, int2 *e1, double *v1, int2 *e2,
                double *v2, hipStream_t st);
int jk_m2_square(int m1, const double *m2, double *sq, hipStream_t st);
int jk_sigma_fast(int prec, int nvec, const FastArgs &a, const void *x, double *y,
                  hipStream_t st);
