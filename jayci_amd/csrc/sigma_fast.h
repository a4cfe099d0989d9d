// Matrix-free sigma from per-string excitation tables: shared declarations.
#pragma once
#include "jci_internal.h"

// device view of one string set's excitation tables (see sigma_fast.hip)
// Entries of a string are grouped by target group: the entries of string I that go to
// group g are [off[I] + goff[I*(G+1) + g], off[I] + goff[I*(G+1) + g + 1]).
struct ExcView {
    const long long *off1, *off2;
    const int2 *e1, *e2;
    const double *v1, *v2;
    const int *goff1, *goff2;
    int G;
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
    int m1int;                   // pairs of internal (doubly occupied + active) orbitals, the
                                 // leading rows and columns of (P|Q)
    int W;                       // beta strings per tile (64, 128, 256 or 512)
    int A;                       // alpha singles per staged chunk (1, 2, 4 or 8; 0 = by LDS size)
    long long ntiles;
    int xg;                      // tiles per XCD run (0 / 1: blockIdx order)
    int cm;                      // tile order in a block: 0 string-major, 1 chunk-major
    int tout;                    // 1: the descriptor is the transpose of the output's (the
                                 // determinant of lane ib_loc, string ia_loc is y[boff +
                                 // ib_loc * gn(p) + ia_loc]; the diagonal is read there too)
    int src_t;                   // 1: reads the block-transposed input xT, 0: x
    long long ldy;               // output layout: 0 = determinant-major [N][K], else vector k
                                 // at y + k*ldy
};

// single-block (FCI) specialisation
#define FCI_MAXSINGLES 256   // alpha singles per string: na * (n - na) <= 256
#define FCI_MAXORB 16        // packed integrals in LDS: n <= 16
struct FciArgs {
    const uint64_t *ma, *mb;
    ExcView ta, tb;
    const double *m2;            // packed (ij|kl)
    const double *hd;
    long long na, nb;            // alpha / beta strings
    int m1, m2len, m2len_pad;
};
int jk_sigma_fci(int prec, int nvec, const FciArgs &a, const void *x, double *y, hipStream_t st);

int jk_exc_count(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                 const unsigned long long *binom, long long *cnt1, long long *cnt2,
                 int *goff1, int *goff2, hipStream_t st);
int jk_exc_fill(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                const unsigned long long *binom, const double *m1, const double *m2,
                const long long *off1, const long long *off2, const int *goff1,
                const int *goff2, int2 *e1, double *v1, int2 *e2, double *v2, hipStream_t st);
int jk_m2_square(int m1, const double *m2, double *sq, hipStream_t st);
int jk_sigma_fast(int prec, int nvec, const FastArgs &a, const void *x, double *y,
                  hipStream_t st);
// the same sigma in two passes: main passes (a: one spin's singles and doubles, the
// diagonal and the alpha-beta products; they write y), then complements (t: the other spin's
// singles and doubles; they add), each launch over its own blocks on the original descriptor
// dD (reading x) or on the transposed one from jci_transpose_desc (spins and tables swapped,
// reading xT, the block-transposed copy of x made here: nvec * N elements of the input's
// type; tt_off / ntt: 16 x 16 transpose tiles per block of dD)
int jk_sigma_fast_split(int prec, int nvec, const SpaceDesc *dD, const FastArgs *a, int na,
                        const FastArgs *t, int nt, const long long *tt_off, long long ntt,
                        const void *x, void *xT, double *y, hipStream_t st);
// one sigma as several fused launches, each over its own blocks (tile-width classes)
int jk_sigma_fast_multi(int prec, int nvec, const FastArgs *a, int na, const void *x, double *y,
                        hipStream_t st);
