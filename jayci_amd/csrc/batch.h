// Batched Davidson: per-instance descriptors shared by batch.cpp and batch_kernels.hip.
#pragma once
#include "jci_internal.h"
#include "sigma_fast.h"
#include "sigma_gemm.h"

// launch header in device memory: the persistent kernels read their work counts here
struct BHeader {
    int n;                 // instances in the descriptor array
    long long total;       // flattened blocks (vector kernels) or tiles (sigma)
    int kmax;              // largest k of the step
};

#define JB_GRID 2048       // persistent work-groups of the vector kernels
#define JB_NCLASS 4        // basis classes of the combine and Gram launches (16/32/48/64 vectors)
#define JB_HCLASS 4        // first device header slot of the class launches (slots 4..7)
#define JB_SGRID 1024      // persistent work-groups of the sigma kernel

// descriptor of one instance for the batched vector kernels (meaning of the fields per
// kernel is given at each kernel)
struct BDesc {
    long long N, ld;
    const double *A, *B;   // inputs
    double *O;             // output vectors
    const double *C;       // small coefficient matrix (m x k, row-major)
    const double *th;      // Ritz values (residual kernel)
    const double *hd;      // diagonal (residual kernel)
    double *out;           // small result (m x k Gram block / k norms)
    int m, k;
    double beta;
    long long blk0;        // first flattened block of the instance
    int nblk;              // blocks of the instance (reductions: rows split evenly)
    long long pofs;        // offset of the instance's partials
};

// descriptor of one instance for the batched sigma
struct BSigma {
    FastArgs a;
    const void *x;         // [N][k] FP64 or FP32
    double *y;             // [N][k]
    int k;
    long long tile0;       // first flattened tile
};

// all launches use fixed grids; n_cap bounds the instance count (reduction grids)
// MAXM-sized Gram launch: mmax bounds every descriptor's m and k (<= JCI_MAXBASIS)
int jb_gram(const BHeader *h, const BDesc *dD, int n_cap, int mmax, double *partial,
            hipStream_t st);
// mmax: the largest m of the launch's descriptors
int jb_combine(const BHeader *h, const BDesc *dD, int mmax, hipStream_t st);
int jb_residual(const BHeader *h, const BDesc *dD, int n_cap, double *partial, hipStream_t st);
int jb_pack(const BHeader *h, const BDesc *dD, int pack, hipStream_t st);
int jb_to32(const BHeader *h, const BDesc *dD, hipStream_t st);
int jb_sigma(int prec, int W, int K, const BHeader *h, const BSigma *dS, int m1max,
             hipStream_t st);

// access to a Hamiltonian's matrix-free arguments (jci.cpp)
int jci_ham_fastargs(jci_ham *h, FastArgs *out);
const double *jci_ham_hd(jci_ham *h);
long long jci_ham_ndets(jci_ham *h);
jci_ctx *jci_ham_ctx(jci_ham *h);
int jci_ham_guess_block(jci_ham *h, int refdim, int nguess, int diag_guess, double *V, long long ld);
int jci_ham_gemmfci(jci_ham *h, GemmFci *out);
