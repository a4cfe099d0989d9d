// Dense-string sigma for full-CI spaces (one alpha group, one beta group, one block).
//
// With X_k the k-th vector viewed as an Na x Nb matrix (alpha string rows, beta string
// columns, determinant index Ia*Nb + Ib), the Hamiltonian splits into
//
//   sigma_k = Haa X_k + X_k Hbb^T + sum_{P,Q} (P|Q) A_P X_k B_Q^T
//
// Haa / Hbb: same-spin string Hamiltonians (one-electron part, same-spin two-electron part
// and the same-spin diagonal), held dense: for the spaces this path takes (n <= 15 orbitals)
// they are 10-30 % filled, so an MFMA product beats a sparse walk.  A_P / B_Q: the
// single-replacement operators E_ij + E_ji of an orbital pair P = tri(i,j) on alpha / beta
// strings, including the diagonal pairs (E_kk = occupation of k), which carry the
// opposite-spin Coulomb terms of the diagonal and of the singles.  The opposite-spin part
// is evaluated as
//
//   D[Q][k,Ja,Ib] = sum_{Jb : <Ib|B_Q|Jb> != 0} sign * X_k[Ja][Jb]        (beta lists)
//   E[P][k,Ja,Ib] = sum_Q (P|Q) D[Q][k,Ja,Ib]                            (MFMA, D in LDS)
//   sigma_k[Ia][Ib] += sum_{(Ja,P,sign) in alpha list of Ia} sign * E[P][k,Ja,Ib]
//
// E[P][.,Ja,.] is read only for the pairs P that connect Ja to some Ia: the pairs of the
// replacement list of Ja itself (a single i -> a of Ia lands on Ja exactly when a -> i
// is a single of Ja; the diagonal pairs are those of Ja's occupied orbitals).  So E is
// kept as La rows per alpha string: row p of string Ja is the pair of entry p of Ja's
// alpha list, and each alpha-list entry carries, in bits 24..30, the position of its pair
// in the list of its target string.
//
// The diagonal pairs of an alpha string all land on the string itself with sign +, so
// their E rows are only ever summed: the alpha list holds one entry for them, pair
// GF_PDIAG, whose row of (P|Q) is sum_{k occupied in Ja} (kk|Q).  La = a(n-a) + 1 instead
// of a(n-a) + a (46 rows instead of 50 for 5 electrons in 14 orbitals: three row tiles of
// 16 instead of four).  Beta lists keep one entry per diagonal pair, since D is needed
// for every Q.
//
// Matrix elements are those of the Slater-Condon rules in slater.h (the corrected
// same-spin doubles of mpi_source/action_util.c:215-240).
#pragma once
#include "jci_internal.h"

#define GF_MAXMP 128          // padded orbital-pair count (n <= 15)
#define GF_MAXSTR 65535       // strings per spin addressable by a 16-bit list entry
#define GF_MAXLIST 64         // replacement-list entries per string: b(n-b+1) <= 64 for n <= 15

// list entry: bits 0..15 target string, 16..23 orbital pair, bit 31 sign; alpha lists also
// bits 24..30: position of the pair in the target string's list (the E row)
#define GF_PDIAG 0xFF         // alpha-list pair: sum of the string's occupied diagonal pairs
struct GemmFci {
    long long na, nb;            // alpha / beta strings
    int m1, MP;                  // orbital pairs, padded to a multiple of 16
    int La, Lb;                  // list entries per alpha / beta string
    const double *H64[2];        // dense same-spin Hamiltonians, [Na][Na] and [Nb][Nb]
    const float *H32[2];
    const double *M64;           // (P|Q), MP x MP, zero padded
    const float *M32;
    const uint32_t *la, *lb;     // [Na][La], [Nb][Lb]
    const double *D64;           // [Na][MP]: the (P|Q) row of each alpha string's GF_PDIAG
    const float *D32;            // entry, sum of (kk|Q) over its occupied orbitals k
};

// One instance's sigma in a grouped launch.  Every kernel of the sigma walks a flattened
// space of work units (FP32 conversion blocks, same-spin product tiles, E column tiles,
// gather blocks); base[] holds the first unit of the task in each space, and the header
// the totals, so one persistent launch per kernel covers any number of instances.
struct GfTask {
    GemmFci g;
    int K;                       // vectors
    const double *x64;           // FP64 input columns, vector k at x64 + k*ldx64
    long long ldx64;
    float *x32;                  // FP32 copies [K][N] (FP32 launches)
    double *y;                   // output columns, vector k at y + k*ldy (overwritten)
    long long ldy;
    void *E;                     // intermediate, La rows of K*N elements, row stride lde
    long long lde;               // K*N rounded up to GF_EALIGN elements (jk_gf_lde)
    long long base[4];           // first unit: conversion, same-spin, E, gather
};
// E rows start on 512-byte boundaries (FP64) / 256-byte (FP32): the tile stores of the E
// kernel then write whole 64-byte sectors
#define GF_EALIGN 64
static inline long long jk_gf_lde(long long ncols) { return (ncols + GF_EALIGN - 1) / GF_EALIGN * GF_EALIGN; }
struct GfHeader {
    int n;                       // tasks
    long long total[4];          // units per kernel
};
#define GF_CV 0
#define GF_SS 1
#define GF_E 2
#define GF_GA 3

// column-tile width of the E kernel (autotuned): FP64 16 or 32, FP32 32 or 64
#define GF_BN64_DEFAULT 32
#define GF_BN32_DEFAULT 32
int jk_gf_bn_valid(int prec, int bn);
// work units of one task per kernel for the given precision and E tile width
void jk_gf_units(int prec, int bn, const GfTask &t, long long units[4]);
// the grouped sigma: (FP32: conversion) -> E -> same-spin product -> gather.  hdr_host
// (may be NULL) sizes the grids to the work; NULL launches the fixed persistent grids
// that read the totals from d_hdr (graph capture).
int jk_gf_group(int prec, int bn, const GfHeader *d_hdr, const GfTask *d_tasks,
                const GfHeader *hdr_host, hipStream_t st);

// build the dense same-spin matrix of one spin from the excitation tables
int jk_gf_dense(long long nstr, const uint64_t *mask, const long long *off1, const int2 *e1,
                const double *v1, const long long *off2, const int2 *e2, const double *v2,
                const double *m1, const double *m2, double *H, hipStream_t st);
// replacement lists of one spin: the singles, then one entry per occupied diagonal pair
// (diag_one = 0, beta) or one GF_PDIAG entry for all of them (diag_one = 1, alpha)
int jk_gf_lists(long long nstr, const uint64_t *mask, const long long *off1, const int2 *e1,
                int L, int diag_one, uint32_t *out, hipStream_t st);
// alpha lists: out = in with bits 24..30 set to the E row of each entry; *missing (device)
// counts entries whose pair is absent from the target's list (0 for a valid list)
int jk_gf_rank(long long nstr, int L, const uint32_t *in, uint32_t *out, int *missing,
               hipStream_t st);
// (P|Q) padded to MP x MP in FP64 and FP32
int jk_gf_m2pad(int m1, int MP, const double *m2sq, double *o64, float *o32, hipStream_t st);
// rows of the GF_PDIAG entries: o[Ja][Q] = sum_{k occupied in Ja} M64[kk][Q], FP64 and FP32
int jk_gf_diagrows(long long nstr, int MP, const uint64_t *mask, const double *M64, double *o64,
                   float *o32, hipStream_t st);
