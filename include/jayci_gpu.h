/*
 * jayci_gpu.h -- C ABI of libjayci_gpu: the MI355X sigma (H v) build and Davidson
 * eigensolver of JayCI.
 *
 * Two layers:
 *   1. Drop-in entry points with the reference's exact names and signatures.  A JayCI
 *      build links them in place of action_util.o, cimapping.o, davidson.o,
 *      citruncate.o and the output writers (reference Makefile link list of `jayci`).
 *   2. A reentrant handle API (jci_*) underneath, for callers that keep several
 *      solves resident on one GPU.
 *
 * All pointers are plain host pointers unless a parameter name ends in _dev (device
 * memory).  Integrals use the reference's packed 1-based layout: moints1 has
 * n(n+1)/2 entries addressed by index1e(i,j)-1 and moints2 m1(m1+1)/2 entries
 * addressed by index2e(i,j,k,l)-1 (source/moindex.c:14-41).  Orbitals 1..n are bits
 * 0..n-1 of a 64-bit string mask (n <= 64).  Determinant order is the reference's:
 * (alpha group, beta group) blocks in citrunc order, alpha-major inside a block.
 *
 * Errors: entry points returning int give 0 on success and a nonzero code otherwise
 * (-1 invalid argument or unsupported case, -2 HIP runtime failure, -3 out of memory,
 * 1 Davidson not converged within maxiter); jci_last_error() describes the last failure
 * of the calling thread, and the reference-named entry points also print it to stderr
 * as "*** Error: <function>: <code>" like source/errorlib.c.
 */
#ifndef JAYCI_GPU_H
#define JAYCI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define JCI_RESTRICT
#else
#define JCI_RESTRICT restrict
#endif

/* ---- reference-format types -------------------------------------------------------- */

/* source/include/binarystr.h:6-11 (byte1: internal-orbital mask; virtx: orbitals above
 * ninto; byte2 unused) */
struct occstr {
    long long int byte1;
    long long int byte2;
    int virtx[2];
    int nvrtx;
};

/* source/include/binarystr.h:14-19 */
struct det {
    struct occstr astr;
    struct occstr bstr;
    int cas;
};

/* source/include/cimapping.h:7-17 */
struct section {
    int first;
    int last;
    struct section *next;
    struct section *prev;
};

struct rowmap {
    int nsec;
    struct section *sec;
};

/* ---- drop-in entry points (reference signatures) ------------------------------------- */

/* replaces source/citruncate.c:123 (citrunc_rtnlist); serial truncation rule; returns a
 * malloc'ed determinant list the caller frees with free() */
struct det *citrunc_rtnlist(int aelec, int belec, int orbs, int nfrzc, int ndocc, int nactv,
                            int nfrzv, int xlvl, int *astr_len, int *bstr_len, int *dtrm_len,
                            int *ci_orbs, int *ci_aelec, int *ci_belec);

/* replaces source/cimapping.c:92 (generate_cimap): fills hmap[0..ndets) (allocated by
 * the caller) with the reference's section lists of connected j >= i, computed by the
 * GPU screen; also keeps the connectivity on the device for compute_hv / dvdalg */
int generate_cimap(struct det *dlist, int ndets, int nactv, struct rowmap *hmap);

/* replaces source/cimapping.c:70 (deallocate_cimap): frees the section lists, hmap
 * itself and the device copy */
int deallocate_cimap(struct rowmap *hmap, int ndets);

/* replaces source/action_util.c:106 (compute_hdgls) */
void compute_hdgls(struct det *dlist, int ndets, double *moints1, double *moints2, int aelec,
                   int belec, double *hdgls, int ninto);

/* replaces source/action_util.c:204 (hmatels); corrected same-spin double rule of
 * mpi_source/action_util.c:237-245 */
double hmatels(struct det deti, struct det detj, double *moints1, double *moints2, int aelec,
               int belec, int nactv);

/* replaces source/action_util.c:121 (compute_hv): c += H v over the map in hmap */
void compute_hv(struct det *dlist, int ndets, double *moints1, double *moints2, int aelec,
                int belec, double *JCI_RESTRICT v, double *JCI_RESTRICT c, int ninto,
                struct rowmap *hmap);

/* replaces source/davidson.c:91 (dvdalg): lowest nroots eigenpairs; civec[r][0..ndets)
 * (rows malloc'ed by the caller), cival[r] = eigenvalue + totfrze.  predr 1: guess from
 * the leading refdim x refdim block; predr 2: read civector.dat in the working
 * directory.  Returns 0, or a nonzero code (see above). */
int dvdalg(struct det *dlist, int ndets, double *moints1, double *moints2, int aelec,
           int belec, double *hdgls, int ninto, double totfrze, int maxiter, int krymin,
           int krymax, int nroots, double restol, struct rowmap *hmap, double **civec,
           double *cival, int predr, int plvl, int norbs, int refdim);

/* replaces source/ioutil.c:184 (print_civectors): writes civector.dat */
int print_civectors(int aelec, int belec, int orbs, double **civec, int ndets, int roots,
                    double *cival);

/* replaces source/ioutil.c:223 (read_civectors): reads civector.dat; 0 ok, 1 no file,
 * 2 header mismatch */
int read_civectors(int aelec, int belec, int orbs, double **civec, int ndets, int roots,
                   double *cival);

/* replaces source/write_wavefunction.c:12 (write_wavefunction): writes ciwvfcn.dat,
 * reading &general from jayci.in */
void write_wavefunction(struct det *detlist, int ndets, int nroots, double **civec,
                        double *cival, int orbs, int ninto, int elec);

/* replaces source/ioutil.c:427 (readmointegrals), which calls the Fortran readmoints
 * (source/readmoints.f90:21): reads the SIFS integral file (8-byte integers, 4-byte record
 * markers, fsplit 1) into moints1 (H1, or the three dipole arrays for itype 3) and
 * moints2, *nuc_rep = energy(1), *fcenergy = frozen-core energy.  The reference opens
 * "moints" and only prints moflname; here a non-empty moflname names the file.  On a bad
 * file it prints the reference's message and ends the program, as the reference's STOP. */
void readmointegrals(double *moints1, double *moints2, int itype, int orbitals, char *moflname,
                     int m1len, int m2len, double *nuc_rep, double *fcenergy);
/* the same reader returning instead of ending the program: 0, or -1 with the message in
 * jci_last_error(); en[0] nuclear repulsion, en[1] frozen-core energy */
int jci_readmoints(const char *path, int type1n, int orbitals, double *moints1, long long m1len,
                   double *moints2, long long m2len, double *en);

/* replace source/ioutil.c:332 (readgeninput) and :280 (readdaiinput): namelists of
 * jayci.in in the working directory (source/readnamelist.f90:38-41, defaults :79-85);
 * restol is parsed at full precision */
void readgeninput(int *elec, int *orbs, int *nfrozen, int *ndocc, int *nactive, int *xlevel,
                  int *nfrzvirt, int *printlvl, int *printwvf, int *err);
void readdaiinput(int *maxiter, int *krymin, int *krymax, int *nroots, int *prediagr,
                  int *refdim, double *restol, int *err);

/* ---- handle API ------------------------------------------------------------------------ */

typedef struct jci_ctx jci_ctx;
typedef struct jci_space jci_space;
typedef struct jci_ham jci_ham;

#define JCI_RULE_SERIAL 0   /* source/citruncate.c:291-312 */
#define JCI_RULE_MRCISD 1   /* mpi_source/citruncate.c:414-471 */

#define JCI_PAIRS_FP64 1    /* keep the connected-pair list with FP64 values */
#define JCI_PAIRS_FP32 2    /* ... and/or with FP32 values */
#define JCI_PAIRS_UPPER 4   /* store only j > i (symmetric sigma with atomics) */
#define JCI_DIRECT_TABLES 8 /* per-string excitation tables for the matrix-free sigma */

#define JCI_SIGMA_AUTO 0
#define JCI_SIGMA_PAIRS64 1
#define JCI_SIGMA_PAIRS32 2
#define JCI_SIGMA_PAIRS_SYM64 3
#define JCI_SIGMA_DIRECT64 4
#define JCI_SIGMA_DIRECT32 5
#define JCI_SIGMA_FAST64 6  /* matrix-free, excitation tables, FP64 */
#define JCI_SIGMA_FAST32 7  /* matrix-free, excitation tables, FP32 accumulation */
#define JCI_SIGMA_FCI64 8   /* matrix-free, single-block (FCI) spaces, n <= 16, FP64 */
#define JCI_SIGMA_FCI32 9   /* ... FP32 accumulation */
#define JCI_SIGMA_GEMM64 10 /* dense-string full-CI sigma on MFMA (n <= 15, <= 65535 strings
                               per spin), FP64 */
#define JCI_SIGMA_GEMM32 11 /* ... FP32 operands and accumulation */

#define JCI_PREC_FP64 0
#define JCI_PREC_MIXED 1

#define JCI_GUESS_BLOCK 1   /* leading refdim block (source/initguess_sbd.c:24) */
#define JCI_GUESS_USER 2    /* caller vectors (source/initguess_roldv.c:20) */
#define JCI_GUESS_DIAG 3    /* block of the refdim lowest diagonal elements */
#define JCI_GUESS_UNIT 4    /* unit vectors on determinants 0..krymin-1 (the MPI driver's
                               prediagr 2, mpi_source/pdavidson.c:418-431) */

#define JCI_RESTART_RITZ 0  /* the lowest krymin Ritz vectors (source/davidson.c:91-288) */
#define JCI_RESTART_GDK 1   /* those and the previous iteration's Ritz vectors of the tracked
                               pairs (GD+k: the restart keeps the direction each pair last
                               moved in) */

#define JCI_MAXROOTS 8

const char *jci_last_error(void);
int jci_device_count(int *n);
/* free and total device memory in bytes (hipMemGetInfo on that device) */
int jci_device_memory(int device, size_t *free_bytes, size_t *total_bytes);

int jci_ctx_create(int device, jci_ctx **out);
void jci_ctx_destroy(jci_ctx *ctx);
int jci_ctx_sync(jci_ctx *ctx);
/* the HIP stream the context's work is queued on (a hipStream_t) */
void *jci_ctx_stream(jci_ctx *ctx);

int jci_space_create(jci_ctx *ctx, int aelec, int belec, int norbs, int ndocc, int nactv,
                     int xlvl, int rule, jci_space **out);
void jci_space_destroy(jci_space *sp);
long long jci_space_ndets(const jci_space *sp);
long long jci_space_nstrings(const jci_space *sp, int spin);
int jci_space_ngroups(const jci_space *sp, int spin);
/* groups as rows (docc, actv, virt, start, count) */
int jci_space_groups(const jci_space *sp, int spin, long long *out);
int jci_space_nblocks(const jci_space *sp);
/* blocks as rows (alpha group, beta group, first determinant) */
int jci_space_blocks(const jci_space *sp, long long *out);
int jci_space_strings(jci_space *sp, int spin, uint64_t *masks);
/* alpha/beta masks of determinants first .. first+count-1 */
int jci_space_dets(jci_space *sp, long long first, long long count, uint64_t *a, uint64_t *b);
/* address of determinant (a, b), or -1 if it is not in the space */
long long jci_space_index(jci_space *sp, uint64_t a, uint64_t b);

int jci_ham_create(jci_space *sp, const double *moints1, const double *moints2, int flags,
                   jci_ham **out);
void jci_ham_destroy(jci_ham *h);
long long jci_ham_npairs(const jci_ham *h);
int jci_ham_hdiag(jci_ham *h, double *out);
/* indices of the nb lowest diagonal elements, ascending by (value, index): the determinants
 * of the JCI_GUESS_DIAG initial block (no counterpart in the reference, whose guess takes
 * the leading refdim determinants, source/initguess_sbd.c:24) */
int jci_ham_lowest_diag(jci_ham *h, int nb, long long *idx);
/* connected off-diagonal entries of the given rows (both triangles, unsorted):
 * row t's entries at [t*cap, t*cap + counts[t]) */
int jci_ham_rows(jci_ham *h, const long long *rows, int nrows, long long cap, long long *cols,
                 double *vals, int *phase, int *deg, long long *counts);
/* excitation-table entries of one string (Hamiltonians built with JCI_DIRECT_TABLES), the
 * per-string singles (kind 1) or doubles (kind 2) to strings of the space, as the matrix-free
 * sigma reads them: targets[] the target string's index in its spin's string list, info[]
 * singles i | f<<6 | negative<<12 | target group<<13 | pair tri(i,f)<<18 (tri(a,b) = a(a+1)/2 + b
 * for a >= b), doubles i0 | i1<<6 | f0<<12 | f1<<18 |
 * negative<<24 | target group<<25 (0-based orbitals), vals[] the one-spin element (singles:
 * unsigned h_if + sum_{k in string, k != i} [(kk|if) - (ki|kf)]; doubles: signed).  No
 * counterpart in the reference (its strings carry yij/cij tables instead,
 * mpi_source/citruncate.c:258-342); for checking the tables.  Returns the entry count (up
 * to cap written) or < 0. */
long long jci_ham_string_excitations(jci_ham *h, int spin, long long string, int kind,
                                     long long cap, long long *targets, int *info, double *vals);
/* y = H x for nvec vectors stored determinant-major ([N][nvec]); _dev: device memory */
int jci_sigma_dev(jci_ham *h, int method, int nvec, const double *x_dev, double *y_dev);
int jci_sigma(jci_ham *h, int method, int nvec, const double *x, double *y);
/* jci_sigma_dev without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, serving loops that forbid hipMalloc).
 * Every buffer beyond x and y (the FP32 copy of x, the dense-string intermediate and its
 * descriptors, vector staging) comes from one library-owned device array of
 * jci_sigma_noalloc_capacity() bytes (128 MiB) per GPU, kept until the process ends; a call
 * needing more returns -1 with the sizes in jci_last_error() (the dense-string sigma then
 * runs its vectors in as many passes as the array allows, one vector of up to about
 * 128 MiB / (8 La) determinants in FP64).  All calls on one GPU share the array, so calls
 * that overlap on one GPU race; results equal jci_sigma_dev's bit for bit. */
int jci_sigma_dev_noalloc(jci_ham *h, int method, int nvec, const double *x_dev, double *y_dev);
long long jci_sigma_noalloc_capacity(void);

typedef struct {
    int maxiter, krymin, krymax, nroots, refdim;
    double restol;
    int precision;        /* JCI_PREC_FP64 or JCI_PREC_MIXED */
    double mixed_switch;  /* residual norm below which FP32 sigma hands over to FP64 (it
                           * also hands over when the largest residual, below 100 times
                           * this or below 1e-2, has made no new 10 % minimum for 6
                           * iterations: the FP32 noise floor) */
    int sigma_fp64, sigma_fp32; /* JCI_SIGMA_* used in each phase (0 = automatic) */
    int guess;            /* JCI_GUESS_* */
    double totfrze;       /* added to the eigenvalues (nuclear repulsion + frozen core) */
    int verbose;
    int nbuffer;          /* extra Ritz pairs expanded every iteration but not required to
                             converge (guards against converging onto a higher root) */
    double etol;          /* energy criterion (0 = off): a requested root also needs
                             |r|^2 < etol * gap, gap = distance of its Ritz value to the
                             nearest Ritz value outside the requested set (|r|^2 / gap bounds
                             the eigenvalue error) */
    double buftol;        /* buffer pairs stop expanding once their residual norm is below
                             buftol (0: they expand until restol, like the requested roots) */
    int restart;          /* JCI_RESTART_*: what a full subspace collapses onto */
} jci_solve_params;

typedef struct {
    int iterations, converged, nbasis;
    long long nsigma64, nsigma32;   /* single-vector sigma applications per precision */
    double rnorm[JCI_MAXROOTS];
    double seconds, sigma_seconds;
} jci_solve_stats;

void jci_solve_defaults(jci_solve_params *p);
/* lowest p->nroots eigenpairs; evecs (may be NULL) receives nroots vectors of N
 * (vector-major); guess (JCI_GUESS_USER) holds krymin vectors of N */
int jci_solve(jci_ham *h, const jci_solve_params *p, const double *guess, double *evals,
              double *evecs, jci_solve_stats *stats);

/* The MPI driver's binary vector files: int32 dim, int32 len, then dim vectors of len
 * doubles (native byte order, no record markers).
 * jci_write_vectors_ufmt replaces mpi_source/mpi_utilities.c:353
 * (print_gavectors2file_dbl_ufmt) for host rows vecs[0..dim-1]: 0 ok, 2 cannot open,
 * 6 write error.  jci_read_vectors_ufmt replaces mpi_source/mpi_utilities.c:426
 * (read_gavectorsfile_dbl_ufmt): 0 ok, 2 cannot open, 3 / 4 short header, 5 the file's
 * dim or len differs from the arguments, 6 short data (the reference's codes 2-5; it
 * reports short data by a nonzero position instead). */
int jci_write_vectors_ufmt(const char *fname, int len, int dim, double *const *vecs);
int jci_read_vectors_ufmt(const char *fname, int len, int dim, double **vecs);

/* replaces the Davidson step of the MPI driver, mpi_source/pdavidson.c:39 (pdavidson),
 * for a handle: prediagr 1 guess from the leading refdim block, 2 unit vectors on the
 * first krymin determinants (krymin <= refdim), 3 krymin vectors read from
 * dir/ci.restart.  Writes dir/civec (the nroots Ritz vectors) and, when not converged,
 * dir/ci.restart (the lowest krymin Ritz vectors) and dir/hvvec (nroots vectors of zeros:
 * the reference clears its sigma block before its loop ends); dir NULL = working
 * directory.  evals (may be NULL): nroots energies + p->totfrze.  Returns 0 converged,
 * 1 not converged (files written; the reference returns 0 there), a file code above,
 * or < 0 on error. */
int jci_pdavidson(jci_ham *h, const jci_solve_params *p, int prediagr, const char *dir,
                  double *evals, jci_solve_stats *stats);

/* G = V^T W for tall-skinny blocks on the FP64 matrix cores (the Davidson projection
 * kernel): V is m vectors of length n, W is w vectors (vector-major, host memory),
 * m, w <= 64; G is m x w row-major */
int jci_gram(jci_ctx *ctx, int m, int w, long long n, const double *V, const double *W,
             double *G);

/* Measurement helpers.  jci_bench_sigma times `reps` back-to-back sigma launches of the
 * given method on nvec random vectors with HIP events recorded on the context stream
 * (the stream the kernel runs on) and returns the mean milliseconds per launch.
 * jci_bench_copy does the same for a plain device copy of `bytes` bytes (bandwidth
 * probe and the byte-counter calibration of a profiler run). */
int jci_bench_sigma(jci_ham *h, int method, int nvec, int reps, double *ms_per_launch);
int jci_bench_copy(jci_ctx *ctx, long long bytes, int reps, double *ms_per_launch);

/* Autotuner: times every sigma variant the Hamiltonian holds (pair list FP64/FP32;
 * matrix-free FP64/FP32 with beta tiles of 64, 128 and 256 strings; the single-block
 * kernel and the dense-string MFMA sigma where they apply) on nvec vectors and
 * makes the fastest FP64 / FP32 ones what JCI_SIGMA_AUTO resolves to.  times (may be
 * NULL) gets ms per launch indexed by JCI_SIGMA_* (0 where not available). */
int jci_ham_autotune(jci_ham *h, int nvec, int reps, int *best64, int *best32, double *times);
/* times[] above has JCI_SIGMA_GEMM32 + 1 entries */
/* beta strings per tile of the matrix-free sigma (0 without excitation tables) */
int jci_ham_fast_width(const jci_ham *h);
/* tile shape of the matrix-free table sigma (the autotuner sets it): *width beta strings per
 * work-group (64, 128 or 256; 0 = per block, the smallest of 64 .. 512 that covers the
 * block's row, one launch per width) and *chunk alpha singles staged together (1, 2, 4 or 8;
 * 0 = the most that fit 64 KB of LDS, at most 4); values >= 0 set, -1 keeps; both are
 * returned */
int jci_ham_fast_tiles(jci_ham *h, int *width, int *chunk);
/* *split 0: the table sigma runs as one fused kernel; 1..3: as two passes over the input
 * and a block-transposed copy of it (N * nvec more device elements): one spin's singles and
 * doubles with the alpha-beta products, then the other spin's singles and doubles.  The
 * lanes of a tile run over the strings of one spin: beta strings for every block (1), alpha
 * strings for the blocks whose beta rows are short (below 256 strings) and alpha columns
 * longer (2), alpha strings for every block (3).  -1 keeps; the value in use is returned
 * (default 2 for spaces of more than one block, else 0; the autotuner sets it) */
int jci_ham_fast_split(jci_ham *h, int *split);
/* tile order of the table sigma.  *group > 1: runs of that many consecutive tiles go to one
 * XCD (they share its L2; the runs go round-robin over the 8 XCDs); 0 or 1: blockIdx order.
 * *chunk_major 1: inside a block, consecutive tiles are consecutive strings on one lane
 * chunk (neighbouring strings share most excitation targets, so such tiles read the same
 * vector rows over the same columns); 0: the chunks of one string are consecutive.  -1
 * keeps; the values in use are returned */
int jci_ham_fast_xcd(jci_ham *h, int *group, int *chunk_major);
/* column-tile widths of the dense-string sigma's opposite-spin product (FP64: 16 or 32,
 * FP32: 32 or 64; the autotuner sets them): a positive *bn sets, then both are returned */
int jci_ham_gemm_tiles(jci_ham *h, int *bn64, int *bn32);

/* dense symmetric eigensolver (cyclic Jacobi, FP64) used for the Davidson subspace:
 * eigenvalues ascending, eigenvector i in vec[i*n + 0..n-1] (largest component positive) */
int jci_jacobi(int n, const double *a, double *eval, double *vec);

/* ---- batched solves (many independent CI instances in lockstep) ---------------------------
 * Every active slot advances one Davidson iteration per jci_batch_step; the sigma of all
 * slots' new vectors is one persistent launch (recorded as a HIP graph per precision and
 * block width and replayed).  Slots are filled between steps (continuous batching) and
 * report done when converged.  Hamiltonians need JCI_DIRECT_TABLES.  max_norbs bounds the
 * orbitals of any instance (LDS sizing). */
typedef struct jci_batch jci_batch;
int jci_batch_create(jci_ctx *ctx, int capacity, int max_norbs, jci_batch **out);
void jci_batch_destroy(jci_batch *b);
int jci_batch_set_graphs(jci_batch *b, int on);
/* device pool for the opposite-spin intermediates of the dense-string sigma (default 24 GiB,
 * allocated at the first such sigma, grown to the largest single vector when smaller);
 * a fuller step runs its sigma in several rounds */
int jci_batch_set_pool(jci_batch *b, long long bytes);
/* the same tile widths for the batch's dense-string sigma (values <= 0 keep the current) */
int jci_batch_set_tiles(jci_batch *b, int bn64, int bn32);
/* the batch's tile widths autotuned: each width timed on the dense-string sigma of h (a
 * full-CI Hamiltonian like the batch's largest instances, nvec vectors, reps launches), the
 * fastest per precision set on the batch and on h and returned */
int jci_batch_tune_tiles(jci_batch *b, jci_ham *h, int nvec, int reps, int *bn64, int *bn32);
/* device bytes a slot of ndets determinants takes (basis, sigma and staging vectors) */
long long jci_batch_bytes(long long ndets, int krymax);
int jci_batch_add(jci_batch *b, jci_ham *h, const jci_solve_params *p, int *slot);
/* the initial block of h's next jci_batch_add with parameters p, computed now (on h's
 * context, from any host thread that uses no other call on that context meanwhile), so that
 * the add only uploads it; a later add with other refdim / krymin / guess settings computes
 * its own (no counterpart in the reference: for serving loops that build requests beside
 * the running batch) */
int jci_ham_prepare_guess(jci_ham *h, const jci_solve_params *p);
int jci_batch_step(jci_batch *b, int *nactive);
int jci_batch_run(jci_batch *b, int max_steps);
/* done: 0 running, 1 converged, 2 stopped (maxiter or no new directions) */
int jci_batch_result(jci_batch *b, int slot, int *done, double *evals, double *evecs,
                     jci_solve_stats *stats);
int jci_batch_release(jci_batch *b, int slot);
/* slots whose solve has ended (converged or failed) and that are not released yet: returns
 * their number and writes up to cap of them to slots (may be NULL), in slot order */
int jci_batch_finished(jci_batch *b, int *slots, int cap);
double jci_batch_sigma_seconds(const jci_batch *b);
/* host wall seconds spent per step phase since creation: [0] sigma, [1] Gram and
 * Rayleigh-Ritz, [2] residuals and decisions, [3] selection / collapse / precision
 * hand-over, [4] Gram-Schmidt and orthonormalisation, [5] append */
int jci_batch_phase_seconds(const jci_batch *b, double *out6);

/* Synthetic integrals: h_ii = -2.0 + 0.25 i (1-based i), h_ij ~ N(0, 0.05^2);
 * (ij|kl) = sum_{P=1..2n} L^P_ij L^P_kl with L^P symmetric, L^P_ii ~ U(0.1, 0.3),
 * L^P_ij ~ N(0, 0.02^2).  SplitMix64 stream from `seed`, Box-Muller normals.
 * E_nuc for these inputs is 10.0 (JCI_SYNTH_ENUC). */
#define JCI_SYNTH_ENUC 10.0
int jci_synth_integrals(int norbs, uint64_t seed, double *moints1, double *moints2);

#ifdef __cplusplus
}
#endif
#endif /* JAYCI_GPU_H */
