// Internal declarations shared by the host code and the HIP kernels of libjayci_gpu.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define JCI_MAXG 16          // electron-occupation groups per spin
#define JCI_MAXB (JCI_MAXG * JCI_MAXG)
#define JCI_MAXBASIS 64      // Davidson subspace limit (krymax)
#define JCI_MAXK 8           // vectors per sigma launch
// mixed precision: the FP32 phase also hands over to FP64 when its largest residual, already
// below jci_mixed_near(switch residual), has not reached JCI_MIXED_GAIN times its best for
// JCI_MIXED_STALL iterations (the FP32 sigma's noise floor lies above the switch residual:
// energies of a few hundred Ha put it near 1e-4; far above the floor a slow stretch of the
// Davidson residual is not a stall)
#define JCI_MIXED_GAIN 0.9
#define JCI_MIXED_STALL 6
static inline double jci_mixed_near(double sw) { return sw * 100.0 > 1e-2 ? sw * 100.0 : 1e-2; }

#define JCI_CHECK(x)                                                                   \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            jci_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return -2;                                                                 \
        }                                                                              \
    } while (0)

void jci_set_error(const char *fmt, ...);

// One spin's string set.  Strings of group g occupy [gstart[g], gstart[g]+gn[g]); inside a
// group the composite rank is (r_docc * nsa + r_actv) * nsv + r_virt, each r the
// lexicographic rank of the sub-combination inside its orbital subspace.
struct SpinDesc {
    int nelec, ngrp;
    int gd[JCI_MAXG], ga[JCI_MAXG], gv[JCI_MAXG];
    long long gstart[JCI_MAXG + 1], gn[JCI_MAXG], nsa[JCI_MAXG], nsv[JCI_MAXG];
};

// Whole determinant space: blocks (p,q) in reference order, alpha-major inside a block.
struct SpaceDesc {
    int norbs, ndocc, nactv, nvirt;
    unsigned long long dmask, amask, vmask, allmask;
    SpinDesc s[2];
    int nblk;
    int blkid[JCI_MAXG][JCI_MAXG];
    int bp[JCI_MAXB], bq[JCI_MAXB];
    long long boff[JCI_MAXB + 1];
    long long ndets;
};

// host-side construction (space.cpp)
int jci_build_desc(int na, int nb, int norbs, int ndocc, int nactv, int xlvl, int rule,
                   SpaceDesc *d);
void jci_binom_table(unsigned long long *tab /* 65*65 */);
// the space with alpha and beta swapped (blocks in the same order and at the same offsets,
// each stored beta-major)
void jci_transpose_desc(const SpaceDesc &D, SpaceDesc *T);

#define JCI_API extern "C" __attribute__((visibility("default")))

// ---- device context (jci.cpp) ------------------------------------------------------------
#define JCI_PARTIAL_BLOCKS 2048
struct jci_ctx {
    int device;
    int refs;                   // spaces holding this context
    int closing;                // destroy requested while spaces were alive
    hipStream_t stream;
    double *d_partial;          // JCI_PARTIAL_BLOCKS * 64 * 64 block partials
    double *d_small;            // 64 * 64 small operands / results
    unsigned long long *d_binom;
    double *h_small;            // pinned mirror of d_small
};

// ---- caller-provided sigma buffers (jci.cpp, the allocation-free entry point) ------------
// A device block of `size` bytes at `base`; jci_scratch_carve hands out 256-byte aligned
// pieces from `used` on and fails with -1 past the end.
struct jci_ham;
struct SigmaScratch {
    unsigned char *base;
    size_t size, used;
};
int jci_scratch_carve(SigmaScratch *s, size_t bytes, void **out, const char *what);
// jci_sigma_dev taking every buffer beyond x and y from s (no device allocation)
int jci_sigma_dev_scratch(jci_ham *h, int method, int nvec, const double *x_dev, double *y_dev,
                          SigmaScratch *s);
int jci_ham_device(const jci_ham *h);
// looks up this device's copy of the allocation-free entry point's array (noalloc.hip);
// called once per device with the device current
int jci_noalloc_setup(int device);

// ---- Davidson over an abstract operator (jci.cpp) -------------------------------------
struct SolveOp {
    void *self;
    jci_ctx *ctx;
    long long N;
    const double *d_hd, *d_m1, *d_m2;
    // y = H x, determinant-major [N][k] device arrays
    int (*sigma)(void *self, int method, int k, const double *x, double *y);
    // optional: y = H x on vector-major columns (V + j*ldv -> S + j*lds), stage: room for k
    // determinant-major vectors the callee may use; returns 1 when the method has no such
    // form and the caller must pack to determinant-major
    int (*sigma_cols)(void *self, int method, int k, const double *V, long long ldv, double *S,
                      long long lds, double *stage);
    // masks of determinants d_idx[0..n) into device arrays oa, ob
    int (*masks)(void *self, long long n, const long long *d_idx, uint64_t *oa, uint64_t *ob);
    // concrete sigma method for a requested one (0 = automatic) and precision
    int (*resolve)(void *self, int method, int want32);
};
#include "../../include/jayci_gpu.h"
int jci_solve_core(SolveOp *op, const jci_solve_params *p, const double *guess, double *evals,
                   double *evecs, jci_solve_stats *stats, int nout);
long long jci_ham_ndets(jci_ham *h);
// jci_solve's operator for h, with d_pre (device, may be NULL) as the preconditioner diagonal
int jci_ham_solve_op(jci_ham *h, const double *d_pre, SolveOp *op);
// GD+k restart block: the lowest kc Ritz vectors of the projected matrix G (row-major m x m;
// eigenvectors Yf, vector j at Yf + j*m) followed by the np previous Ritz vectors Yp (vector j
// at Yp + j*mp, mp <= m rows, zero below), each made orthonormal to the columns before it by
// two Gram-Schmidt passes (one left with norm below 1e-8 is dropped), at most kmax columns.
// C (m x kmax) receives the columns row-major (element (i, j) at C[i*cols + j]) and Gc
// (kmax x kmax) C^T G C row-major (cols x cols).  Returns cols.
int jci_gdk_restart(int m, const double *G, const double *Yf, int kc, int mp, int np,
                    const double *Yp, int kmax, double *C, double *Gc);
// subspace sizes of a solve: tracked pairs (roots + buffer pairs, space permitting), the
// restart size krymin (also the initial block) and krymax, for N determinants
void jci_subspace_sizes(const jci_solve_params *p, long long N, int *ntrack, int *krymin,
                        int *krymax);
// a Hamiltonian's prepared initial block (jci_ham_prepare_guess) for these settings: 1 and its
// size and host arrays (U: vector j at U + j*nb), or 0
int jci_ham_prepared_guess(jci_ham *h, int refdim, int nguess, int diag_guess, int *nb,
                           const long long **idx, const double **U);
// V[:, 0..nguess) = the block's vectors scattered to their determinants, on stream st; d_scr
// holds (8 + 8 nguess) nb bytes
int jci_guess_apply(int nb, const long long *idx, const double *U, int nguess, double *V,
                    long long ld, void *d_scr, hipStream_t st);
// jci_solve returning the lowest max(nroots, nout) Ritz vectors
int jci_solve_nout(jci_ham *h, const jci_solve_params *p, const double *guess, double *evals,
                   double *evecs, jci_solve_stats *stats, int nout);

// ---- kernel launchers (kernels.hip) ------------------------------------------------
int jk_gen_strings(const SpaceDesc *d_desc, const SpaceDesc &h, int spin,
                   const unsigned long long *d_binom, uint64_t *d_mask, hipStream_t st);
int jk_hdiag(const SpaceDesc *d_desc, long long ndets, const uint64_t *ma, const uint64_t *mb,
             const double *m1, const double *m2, double *hd, hipStream_t st);
int jk_row_count(const SpaceDesc *d_desc, long long ndets, const uint64_t *ma,
                 const uint64_t *mb, const unsigned long long *binom, long long *cnt,
                 int upper, hipStream_t st);
int jk_row_fill(const SpaceDesc *d_desc, long long ndets, const uint64_t *ma, const uint64_t *mb,
                const unsigned long long *binom, const double *m1, const double *m2,
                const long long *rowptr, int *col, double *val64, float *val32, int upper,
                hipStream_t st);
int jk_rows_debug(const SpaceDesc *d_desc, const uint64_t *ma, const uint64_t *mb,
                  const unsigned long long *binom, const double *m1, const double *m2,
                  const long long *rows, int nrows, long long cap, long long *cols, double *vals,
                  int *phase, int *deg, long long *counts, hipStream_t st);
int jk_scan(const long long *cnt, long long *rowptr, long long n, hipStream_t st);
int jk_spmv(int prec, int nvec, long long N, const long long *rowptr, const int *col,
            const void *val, const double *hd, const void *x, double *y, hipStream_t st);
int jk_spmv_sym(int nvec, long long N, const long long *rowptr, const int *col,
                const double *val, const double *hd, const double *x, double *y, hipStream_t st);
int jk_sigma_generic(int prec, int nvec, const SpaceDesc *d_desc, long long ndets,
                     const uint64_t *ma, const uint64_t *mb, const unsigned long long *binom,
                     const double *m1, const double *m2, const double *hd, const void *x,
                     double *y, hipStream_t st);
int jk_to_f32(const double *x, float *y, long long n, hipStream_t st);

// Davidson vector kernels
int jk_gram(int m, int w, long long N, const double *V, long long ldv, const double *W,
            long long ldw, double *partial, int nblocks, double *out_dev, hipStream_t st);
int jk_combine(int m, int k, long long N, const double *V, long long ldv, const double *Y,
               int ldy, double *out, long long ldo, double beta, hipStream_t st);
int jk_residual(int m, int k, long long N, const double *V, const double *S, long long ld,
                const double *Y, const double *theta, const double *hd, double *T,
                double *partial, int nblocks, double *norms_dev, hipStream_t st);
int jk_pack(int k, long long N, const double *src, long long lds, double *dst, hipStream_t st);
int jk_unpack(int k, long long N, const double *src, double *dst, long long ldd, hipStream_t st);
int jk_colnorms(int k, long long N, const double *X, long long ld, double *partial,
                int nblocks, double *out_dev, hipStream_t st);
int jk_scale_cols(int k, long long N, double *X, long long ld, const double *s, hipStream_t st);
int jk_dense_block(int nb, const uint64_t *ma_rows, const uint64_t *mb_rows, int norbs,
                   const double *m1, const double *m2, double *H, hipStream_t st);
// the nb (<= JCI_GUESS_CHUNK) lowest diagonal elements of each chunk of JCI_GUESS_CHUNK
// determinants: values cv and indices ci, chunk c at [c*nb, c*nb + nb), ascending (value,
// index); entries past N come as (+inf, LLONG_MAX)
#define JCI_GUESS_CHUNK 2048
int jk_lowest_chunks(long long N, const double *hd, int nb, double *cv, long long *ci,
                     hipStream_t st);

// drop-in path on explicit determinant lists
int jk_list_hdiag(long long n, const uint64_t *ma, const uint64_t *mb, int norbs,
                  const double *m1, const double *m2, double *hd, hipStream_t st);
int jk_list_screen_count(long long n, const uint64_t *ma, const uint64_t *mb, long long *cnt,
                         hipStream_t st);
int jk_list_screen_fill(long long n, const uint64_t *ma, const uint64_t *mb,
                        const long long *rowptr, int *col, hipStream_t st);
int jk_list_values(long long nnz_rows, long long n, const uint64_t *ma, const uint64_t *mb,
                   int norbs, const double *m1, const double *m2, const long long *rowptr,
                   const int *col, double *val, hipStream_t st);
