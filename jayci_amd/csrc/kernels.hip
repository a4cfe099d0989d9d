// HIP kernels of libjayci_gpu (gfx950).
//
//   string tables ......... k_gen_strings      (replaces source/straddress.c:28, str_adr2str)
//   diagonal .............. k_hdiag            (source/action_util.c:106, compute_hdgls)
//   row enumeration ....... visit_row          (string-driven rows, mpi_source/pdavidson.c:1964-2067)
//   pair list ............. k_row_count / k_row_fill (replaces source/cimapping.c:92, generate_cimap,
//                           with H values attached)
//   sigma ................. k_spmv / k_spmv_sym / k_sigma_generic (source/action_util.c:121, compute_hv)
//   Davidson vectors ...... k_gram (MFMA f64), k_combine, k_residual, k_colnorms
//   determinant lists ..... k_list_* (the O(N^2) screen of source/cimapping.c:92 on explicit lists)
#include "jci_internal.h"
#include "slater.h"

#define LAUNCH_CHECK()                                                   \
    do {                                                                 \
        hipError_t e_ = hipGetLastError();                               \
        if (e_ != hipSuccess) {                                          \
            jci_set_error("%s:%d launch: %s", __FILE__, __LINE__,        \
                          hipGetErrorString(e_));                        \
            return -2;                                                   \
        }                                                                \
    } while (0)

static inline unsigned grid_for(long long n, int per_block)
{
    long long g = (n + per_block - 1) / per_block;
    return (unsigned)(g > 0 ? g : 1);
}

// ---------------------------------------------------------------------------------------
// space helpers

__device__ __forceinline__ int find_block(const SpaceDesc *d, long long I)
{
    int lo = 0, hi = d->nblk - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d->boff[mid] <= I) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t shr(uint64_t x, int s) { return s >= 64 ? 0ULL : (x >> s); }

// address of string mask m of spin s (or -1); group in *g
__device__ __forceinline__ long long str_index(const SpaceDesc *d, int s, uint64_t m,
                                               const unsigned long long *binom, int *g)
{
    int nd = jci_popc(m & d->dmask), na = jci_popc(m & d->amask), nv = jci_popc(m & d->vmask);
    const SpinDesc &S = d->s[s];
    int gg = -1;
    for (int t = 0; t < S.ngrp; t++)
        if (S.gd[t] == nd && S.ga[t] == na && S.gv[t] == nv) { gg = t; break; }
    if (gg < 0) return -1;
    long long rd = jci_lex_rank(m & d->dmask, d->ndocc, nd, binom);
    long long ra = jci_lex_rank(shr(m & d->amask, d->ndocc), d->nactv, na, binom);
    long long rv = jci_lex_rank(shr(m & d->vmask, d->ndocc + d->nactv), d->nvirt, nv, binom);
    *g = gg;
    return S.gstart[gg] + (rd * S.nsa[gg] + ra) * S.nsv[gg] + rv;
}

// ---------------------------------------------------------------------------------------
// string tables: one thread per string, unranking its composite rank

__global__ void k_gen_strings(const SpaceDesc *__restrict__ d, int spin,
                              const unsigned long long *__restrict__ binom,
                              uint64_t *__restrict__ out)
{
    const SpinDesc &S = d->s[spin];
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.gstart[S.ngrp]) return;
    int g = 0;
    while (g + 1 < S.ngrp && S.gstart[g + 1] <= i) g++;
    long long r = i - S.gstart[g];
    long long nav = S.nsa[g] * S.nsv[g];
    long long rd = r / nav, ra = (r / S.nsv[g]) % S.nsa[g], rv = r % S.nsv[g];
    uint64_t m = jci_lex_unrank(rd, d->ndocc, S.gd[g], binom);
    if (S.ga[g] > 0) m |= jci_lex_unrank(ra, d->nactv, S.ga[g], binom) << d->ndocc;
    if (S.gv[g] > 0) m |= jci_lex_unrank(rv, d->nvirt, S.gv[g], binom) << (d->ndocc + d->nactv);
    out[i] = m;
}

int jk_gen_strings(const SpaceDesc *d_desc, const SpaceDesc &h, int spin,
                   const unsigned long long *d_binom, uint64_t *d_mask, hipStream_t st)
{
    long long n = h.s[spin].gstart[h.s[spin].ngrp];
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gen_strings, dim3(grid_for(n, 256)), dim3(256), 0, st, d_desc, spin,
                       d_binom, d_mask);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// determinant decode

struct DetRef {
    long long ia, ib, ia_loc;
    int p, q;
};

__device__ __forceinline__ DetRef decode(const SpaceDesc *d, long long I)
{
    DetRef r;
    int blk = find_block(d, I);
    r.p = d->bp[blk];
    r.q = d->bq[blk];
    long long off = I - d->boff[blk];
    long long nq = d->s[1].gn[r.q];
    r.ia_loc = off / nq;
    r.ia = d->s[0].gstart[r.p] + r.ia_loc;
    r.ib = d->s[1].gstart[r.q] + (off - r.ia_loc * nq);
    return r;
}

__global__ void k_hdiag(const SpaceDesc *__restrict__ d, long long N,
                        const uint64_t *__restrict__ MA, const uint64_t *__restrict__ MB,
                        const double *__restrict__ m1, const double *__restrict__ m2,
                        double *__restrict__ hd)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    DetRef r = decode(d, I);
    hd[I] = jci_diag(MA[r.ia], MB[r.ib], m1, m2);
}

int jk_hdiag(const SpaceDesc *d_desc, long long ndets, const uint64_t *ma, const uint64_t *mb,
             const double *m1, const double *m2, double *hd, hipStream_t st)
{
    hipLaunchKernelGGL(k_hdiag, dim3(grid_for(ndets, 256)), dim3(256), 0, st, d_desc, ndets,
                       ma, mb, m1, m2, hd);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// Row enumeration: every determinant J connected to row I (J != I), with its matrix
// element, sign and excitation degree.  Singles and doubles are generated per spin from
// the string masks; target strings are addressed by rank; blocks not in the space are
// skipped.  Order: beta singles, beta doubles, then per alpha single its own element and
// its products with beta singles, then alpha doubles.

template <bool VAL, class Vis>
__device__ void visit_row(const SpaceDesc *__restrict__ d, const uint64_t *__restrict__ MA,
                          const uint64_t *__restrict__ MB,
                          const unsigned long long *__restrict__ binom,
                          const double *__restrict__ m1, const double *__restrict__ m2,
                          long long I, Vis &vis)
{
    DetRef r = decode(d, I);
    const uint64_t ma = MA[r.ia], mb = MB[r.ib], all = d->allmask;
    const SpinDesc &A = d->s[0], &B = d->s[1];
    // beta singles
    for (uint64_t x = mb; x; x &= x - 1) {
        int i = jci_ctz(x);
        for (uint64_t y = all & ~mb; y; y &= y - 1) {
            int f = jci_ctz(y);
            int gq;
            long long jb = str_index(d, 1, mb ^ (1ULL << i) ^ (1ULL << f), binom, &gq);
            if (jb < 0) continue;
            int bk = d->blkid[r.p][gq];
            if (bk < 0) continue;
            long long J = d->boff[bk] + r.ia_loc * B.gn[gq] + (jb - B.gstart[gq]);
            int sg = jci_sign(mb, i, f);
            vis(J, VAL ? sg * jci_single_bracket(mb, ma, i, f, m1, m2) : 0.0, sg, 1);
        }
    }
    // beta doubles
    for (uint64_t x0 = mb; x0; x0 &= x0 - 1) {
        int i0 = jci_ctz(x0);
        for (uint64_t x1 = x0 & (x0 - 1); x1; x1 &= x1 - 1) {
            int i1 = jci_ctz(x1);
            for (uint64_t y0 = all & ~mb; y0; y0 &= y0 - 1) {
                int f0 = jci_ctz(y0);
                for (uint64_t y1 = y0 & (y0 - 1); y1; y1 &= y1 - 1) {
                    int f1 = jci_ctz(y1);
                    uint64_t jm = mb ^ (1ULL << i0) ^ (1ULL << i1) ^ (1ULL << f0) ^ (1ULL << f1);
                    int gq;
                    long long jb = str_index(d, 1, jm, binom, &gq);
                    if (jb < 0) continue;
                    int bk = d->blkid[r.p][gq];
                    if (bk < 0) continue;
                    long long J = d->boff[bk] + r.ia_loc * B.gn[gq] + (jb - B.gstart[gq]);
                    int ph = 0;
                    double v = VAL ? jci_double_same(mb, i0, i1, f0, f1, m2, &ph) : 0.0;
                    vis(J, v, ph, 2);
                }
            }
        }
    }
    // alpha singles, alone and times beta singles
    for (uint64_t x = ma; x; x &= x - 1) {
        int i = jci_ctz(x);
        for (uint64_t y = all & ~ma; y; y &= y - 1) {
            int f = jci_ctz(y);
            int gp;
            long long ja = str_index(d, 0, ma ^ (1ULL << i) ^ (1ULL << f), binom, &gp);
            if (ja < 0) continue;
            long long ja_loc = ja - A.gstart[gp];
            int sa = jci_sign(ma, i, f);
            int bk = d->blkid[gp][r.q];
            if (bk >= 0) {
                long long J = d->boff[bk] + ja_loc * B.gn[r.q] + (r.ib - B.gstart[r.q]);
                vis(J, VAL ? sa * jci_single_bracket(ma, mb, i, f, m1, m2) : 0.0, sa, 1);
            }
            for (uint64_t xb = mb; xb; xb &= xb - 1) {
                int k = jci_ctz(xb);
                for (uint64_t yb = all & ~mb; yb; yb &= yb - 1) {
                    int l = jci_ctz(yb);
                    int gq;
                    long long jb = str_index(d, 1, mb ^ (1ULL << k) ^ (1ULL << l), binom, &gq);
                    if (jb < 0) continue;
                    int bk2 = d->blkid[gp][gq];
                    if (bk2 < 0) continue;
                    long long J = d->boff[bk2] + ja_loc * B.gn[gq] + (jb - B.gstart[gq]);
                    int sg = sa * jci_sign(mb, k, l);
                    vis(J, VAL ? sg * jci_eri(m2, i, f, k, l) : 0.0, sg, 2);
                }
            }
        }
    }
    // alpha doubles
    for (uint64_t x0 = ma; x0; x0 &= x0 - 1) {
        int i0 = jci_ctz(x0);
        for (uint64_t x1 = x0 & (x0 - 1); x1; x1 &= x1 - 1) {
            int i1 = jci_ctz(x1);
            for (uint64_t y0 = all & ~ma; y0; y0 &= y0 - 1) {
                int f0 = jci_ctz(y0);
                for (uint64_t y1 = y0 & (y0 - 1); y1; y1 &= y1 - 1) {
                    int f1 = jci_ctz(y1);
                    uint64_t jm = ma ^ (1ULL << i0) ^ (1ULL << i1) ^ (1ULL << f0) ^ (1ULL << f1);
                    int gp;
                    long long ja = str_index(d, 0, jm, binom, &gp);
                    if (ja < 0) continue;
                    int bk = d->blkid[gp][r.q];
                    if (bk < 0) continue;
                    long long J = d->boff[bk] + (ja - A.gstart[gp]) * B.gn[r.q] + (r.ib - B.gstart[r.q]);
                    int ph = 0;
                    double v = VAL ? jci_double_same(ma, i0, i1, f0, f1, m2, &ph) : 0.0;
                    vis(J, v, ph, 2);
                }
            }
        }
    }
}

struct VisCount {
    long long n, I;
    int upper;
    __device__ void operator()(long long J, double, int, int) { n += (!upper || J > I); }
};

struct VisFill {
    int *col;
    double *v64;
    float *v32;
    long long pos, I;
    int upper;
    __device__ void operator()(long long J, double v, int, int)
    {
        if (upper && J <= I) return;
        col[pos] = (int)J;
        if (v64) v64[pos] = v;
        if (v32) v32[pos] = (float)v;
        pos++;
    }
};

struct VisDebug {
    long long *cols;
    double *vals;
    int *ph, *dg;
    long long n, cap;
    __device__ void operator()(long long J, double v, int p, int g)
    {
        if (n < cap) { cols[n] = J; vals[n] = v; ph[n] = p; dg[n] = g; }
        n++;
    }
};

template <typename AT, typename XT, int K>
struct VisSigma {
    const XT *x;
    AT acc[K];
    __device__ void operator()(long long J, double v, int, int)
    {
        AT vv = (AT)v;
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += vv * (AT)x[J * K + k];
    }
};

__global__ void k_row_count(const SpaceDesc *__restrict__ d, long long N,
                            const uint64_t *__restrict__ MA, const uint64_t *__restrict__ MB,
                            const unsigned long long *__restrict__ binom,
                            long long *__restrict__ cnt, int upper)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    VisCount v{0, I, upper};
    visit_row<false>(d, MA, MB, binom, nullptr, nullptr, I, v);
    cnt[I] = v.n;
}

int jk_row_count(const SpaceDesc *d_desc, long long ndets, const uint64_t *ma,
                 const uint64_t *mb, const unsigned long long *binom, long long *cnt,
                 int upper, hipStream_t st)
{
    hipLaunchKernelGGL(k_row_count, dim3(grid_for(ndets, 128)), dim3(128), 0, st, d_desc,
                       ndets, ma, mb, binom, cnt, upper);
    LAUNCH_CHECK();
    return 0;
}

__global__ void k_row_fill(const SpaceDesc *__restrict__ d, long long N,
                           const uint64_t *__restrict__ MA, const uint64_t *__restrict__ MB,
                           const unsigned long long *__restrict__ binom,
                           const double *__restrict__ m1, const double *__restrict__ m2,
                           const long long *__restrict__ rowptr, int *__restrict__ col,
                           double *__restrict__ v64, float *__restrict__ v32, int upper)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    VisFill v{col, v64, v32, rowptr[I], I, upper};
    visit_row<true>(d, MA, MB, binom, m1, m2, I, v);
}

int jk_row_fill(const SpaceDesc *d_desc, long long ndets, const uint64_t *ma, const uint64_t *mb,
                const unsigned long long *binom, const double *m1, const double *m2,
                const long long *rowptr, int *col, double *val64, float *val32, int upper,
                hipStream_t st)
{
    hipLaunchKernelGGL(k_row_fill, dim3(grid_for(ndets, 128)), dim3(128), 0, st, d_desc, ndets,
                       ma, mb, binom, m1, m2, rowptr, col, val64, val32, upper);
    LAUNCH_CHECK();
    return 0;
}

__global__ void k_rows_debug(const SpaceDesc *__restrict__ d, const uint64_t *__restrict__ MA,
                             const uint64_t *__restrict__ MB,
                             const unsigned long long *__restrict__ binom,
                             const double *__restrict__ m1, const double *__restrict__ m2,
                             const long long *__restrict__ rows, int nrows, long long cap,
                             long long *cols, double *vals, int *ph, int *dg, long long *counts)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    VisDebug v{cols + (long long)t * cap, vals + (long long)t * cap, ph + (long long)t * cap,
               dg + (long long)t * cap, 0, cap};
    visit_row<true>(d, MA, MB, binom, m1, m2, rows[t], v);
    counts[t] = v.n;
}

int jk_rows_debug(const SpaceDesc *d_desc, const uint64_t *ma, const uint64_t *mb,
                  const unsigned long long *binom, const double *m1, const double *m2,
                  const long long *rows, int nrows, long long cap, long long *cols, double *vals,
                  int *phase, int *deg, long long *counts, hipStream_t st)
{
    hipLaunchKernelGGL(k_rows_debug, dim3(grid_for(nrows, 64)), dim3(64), 0, st, d_desc, ma, mb,
                       binom, m1, m2, rows, nrows, cap, cols, vals, phase, deg, counts);
    LAUNCH_CHECK();
    return 0;
}

// exclusive scan of row counts into rowptr[0..n] (single work-group, chunked; the
// pair-list build is a one-off)
__global__ void k_scan(const long long *__restrict__ cnt, long long *__restrict__ rowptr,
                       long long n)
{
    __shared__ long long part[1024];
    int t = threadIdx.x;
    long long chunk = (n + 1023) / 1024;
    long long b = t * chunk, e = b + chunk < n ? b + chunk : n;
    long long s = 0;
    for (long long i = b; i < e; i++) s += cnt[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        long long run = 0;
        for (int k = 0; k < 1024; k++) { long long v = part[k]; part[k] = run; run += v; }
        rowptr[n] = run;
    }
    __syncthreads();
    long long run = part[t];
    for (long long i = b; i < e; i++) { rowptr[i] = run; run += cnt[i]; }
}

int jk_scan(const long long *cnt, long long *rowptr, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cnt, rowptr, n);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// sigma from the pair list: one 64-lane wavefront per row, coalesced column/value
// stream, gathers of x from cache, wavefront reduction in a fixed order.  x and y are
// determinant-major with K vectors interleaved ([N][K]); y = diag*x + offdiag*x.

template <typename VT, typename AT, typename XT, int K>
__global__ __launch_bounds__(256) void k_spmv(long long N, const long long *__restrict__ rowptr,
                                              const int *__restrict__ col,
                                              const VT *__restrict__ val,
                                              const double *__restrict__ hd,
                                              const XT *__restrict__ x, double *__restrict__ y)
{
    long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= N) return;
    long long b = rowptr[row], e = rowptr[row + 1];
    AT acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = (AT)0;
    for (long long t = b + lane; t < e; t += 64) {
        int j = __builtin_nontemporal_load(col + t);
        AT v = (AT)__builtin_nontemporal_load(val + t);
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += v * (AT)x[(long long)j * K + k];
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 64);
    if (lane == 0) {
        AT dg = (AT)hd[row];
#pragma unroll
        for (int k = 0; k < K; k++) y[row * K + k] = (double)(acc[k] + dg * (AT)x[row * K + k]);
    }
}

template <typename VT, typename AT, typename XT>
static int spmv_k(int nvec, long long N, const long long *rowptr, const int *col, const VT *val,
                  const double *hd, const XT *x, double *y, hipStream_t st)
{
    dim3 g(grid_for(N, 4)), b(256);
    switch (nvec) {
    case 1: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 1>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 2: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 2>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 3: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 3>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 4: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 4>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 5: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 5>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 6: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 6>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 7: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 7>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    case 8: hipLaunchKernelGGL((k_spmv<VT, AT, XT, 8>), g, b, 0, st, N, rowptr, col, val, hd, x, y); break;
    default: jci_set_error("spmv: nvec must be 1..8"); return -1;
    }
    LAUNCH_CHECK();
    return 0;
}

int jk_spmv(int prec, int nvec, long long N, const long long *rowptr, const int *col,
            const void *val, const double *hd, const void *x, double *y, hipStream_t st)
{
    if (prec == 64)
        return spmv_k<double, double, double>(nvec, N, rowptr, col, (const double *)val, hd,
                                              (const double *)x, y, st);
    return spmv_k<float, float, float>(nvec, N, rowptr, col, (const float *)val, hd,
                                       (const float *)x, y, st);
}

// symmetric variant: upper-triangle pairs only; the transposed contribution is scattered
// with hardware FP64 atomics.  y must hold diag*x on entry.
__global__ __launch_bounds__(256) void k_spmv_sym(long long N, const long long *__restrict__ rowptr,
                                                  const int *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y)
{
    long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= N) return;
    long long b = rowptr[row], e = rowptr[row + 1];
    double xi = x[row], acc = 0.0;
    for (long long t = b + lane; t < e; t += 64) {
        int j = __builtin_nontemporal_load(col + t);
        double v = __builtin_nontemporal_load(val + t);
        acc += v * x[j];
        unsafeAtomicAdd(y + j, v * xi);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) unsafeAtomicAdd(y + row, acc);
}

__global__ void k_diag_mul(long long N, const double *__restrict__ hd,
                           const double *__restrict__ x, double *__restrict__ y)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) y[i] = hd[i] * x[i];
}

int jk_spmv_sym(int nvec, long long N, const long long *rowptr, const int *col,
                const double *val, const double *hd, const double *x, double *y, hipStream_t st)
{
    if (nvec != 1) { jci_set_error("spmv_sym: nvec must be 1"); return -1; }
    hipLaunchKernelGGL(k_diag_mul, dim3(grid_for(N, 256)), dim3(256), 0, st, N, hd, x, y);
    hipLaunchKernelGGL(k_spmv_sym, dim3(grid_for(N, 4)), dim3(256), 0, st, N, rowptr, col, val,
                       x, y);
    LAUNCH_CHECK();
    return 0;
}

// matrix-free sigma: one thread per row, rows enumerated on the fly (general spaces)
template <typename AT, typename XT, int K>
__global__ __launch_bounds__(128) void k_sigma_generic(const SpaceDesc *__restrict__ d, long long N,
                                                       const uint64_t *__restrict__ MA,
                                                       const uint64_t *__restrict__ MB,
                                                       const unsigned long long *__restrict__ binom,
                                                       const double *__restrict__ m1,
                                                       const double *__restrict__ m2,
                                                       const double *__restrict__ hd,
                                                       const XT *__restrict__ x,
                                                       double *__restrict__ y)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    VisSigma<AT, XT, K> v;
    v.x = x;
#pragma unroll
    for (int k = 0; k < K; k++) v.acc[k] = (AT)0;
    visit_row<true>(d, MA, MB, binom, m1, m2, I, v);
    AT dg = (AT)hd[I];
#pragma unroll
    for (int k = 0; k < K; k++) y[I * K + k] = (double)(v.acc[k] + dg * (AT)x[I * K + k]);
}

template <typename AT, typename XT>
static int sigma_generic_k(int nvec, const SpaceDesc *d, long long N, const uint64_t *ma,
                           const uint64_t *mb, const unsigned long long *binom, const double *m1,
                           const double *m2, const double *hd, const XT *x, double *y,
                           hipStream_t st)
{
    dim3 g(grid_for(N, 128)), b(128);
    switch (nvec) {
    case 1: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 1>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 2: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 2>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 3: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 3>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 4: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 4>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 5: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 5>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 6: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 6>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 7: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 7>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    case 8: hipLaunchKernelGGL((k_sigma_generic<AT, XT, 8>), g, b, 0, st, d, N, ma, mb, binom, m1, m2, hd, x, y); break;
    default: jci_set_error("sigma: nvec must be 1..8"); return -1;
    }
    LAUNCH_CHECK();
    return 0;
}

int jk_sigma_generic(int prec, int nvec, const SpaceDesc *d_desc, long long ndets,
                     const uint64_t *ma, const uint64_t *mb, const unsigned long long *binom,
                     const double *m1, const double *m2, const double *hd, const void *x,
                     double *y, hipStream_t st)
{
    if (prec == 64)
        return sigma_generic_k<double, double>(nvec, d_desc, ndets, ma, mb, binom, m1, m2, hd,
                                               (const double *)x, y, st);
    return sigma_generic_k<float, float>(nvec, d_desc, ndets, ma, mb, binom, m1, m2, hd,
                                         (const float *)x, y, st);
}

__global__ void k_to_f32(const double *__restrict__ x, float *__restrict__ y, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

int jk_to_f32(const double *x, float *y, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256)), dim3(256), 0, st, x, y, n);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// Davidson vector kernels.  Basis vectors are stored one after another (leading
// dimension ld >= N).

// G = V^T W on MFMA (v_mfma_f64_16x16x4f64).  Each work-group stages 64-row tiles of V and
// W in LDS (coalesced), each wavefront owns 16x16 output tiles and accumulates over the
// row tiles of its work-group; per-work-group partials are reduced in a fixed order by
// k_gram_reduce, so the result is deterministic.
//
// Operand layout of the 16x16x4 f64 MFMA (lane l, 64 lanes):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, register v holds D[l / 16 + 4 * v][l % 16]
#define GRAM_ROWS 32
__global__ __launch_bounds__(256) void k_gram(int m, int w, long long N,
                                              const double *__restrict__ V, long long ldv,
                                              const double *__restrict__ W, long long ldw,
                                              double *__restrict__ partial, long long chunk)
{
    __shared__ double sV[GRAM_ROWS][JCI_MAXBASIS + 1];
    __shared__ double sW[GRAM_ROWS][JCI_MAXBASIS + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mt = (m + 15) / 16, wt = (w + 15) / 16, ntile = mt * wt;
    typedef double v4d __attribute__((ext_vector_type(4)));
    v4d acc[4];
    for (int t = 0; t < 4; t++) acc[t] = (v4d){0.0, 0.0, 0.0, 0.0};
    long long r0 = (long long)blockIdx.x * chunk;
    long long r1 = r0 + chunk < N ? r0 + chunk : N;
    // stage: element (vector c, row r) -> sV[r][c], zero-padded to multiples of 16; element
    // idx = tid + 256 q of the chunk is column idx / GRAM_ROWS, row idx % GRAM_ROWS.  The
    // values are loaded into registers one chunk ahead, so the loads of the next chunk are
    // in flight during the products of this one.
    constexpr int Q = GRAM_ROWS * JCI_MAXBASIS / 256;
    double rv[Q], rw[Q];
    auto gload = [&](long long base) {
        const int rows = (int)((r1 - base) < GRAM_ROWS ? (r1 - base) : GRAM_ROWS);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int idx = tid + 256 * q, c = idx / GRAM_ROWS, r = idx % GRAM_ROWS;
            rv[q] = (idx < GRAM_ROWS * mt * 16 && c < m && r < rows) ? V[(long long)c * ldv + base + r] : 0.0;
            rw[q] = (idx < GRAM_ROWS * wt * 16 && c < w && r < rows) ? W[(long long)c * ldw + base + r] : 0.0;
        }
    };
    if (r0 < r1) gload(r0);
    for (long long base = r0; base < r1; base += GRAM_ROWS) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int idx = tid + 256 * q;
            if (idx < GRAM_ROWS * mt * 16) sV[idx % GRAM_ROWS][idx / GRAM_ROWS] = rv[q];
            if (idx < GRAM_ROWS * wt * 16) sW[idx % GRAM_ROWS][idx / GRAM_ROWS] = rw[q];
        }
        __syncthreads();
        if (base + GRAM_ROWS < r1) gload(base + GRAM_ROWS);
        for (int t = 0; t < 4; t++) {
            int tile = wave + 4 * t;
            if (tile >= ntile) break;
            int ti = tile / wt, tj = tile % wt;
            for (int k0 = 0; k0 < GRAM_ROWS; k0 += 4) {
                double a = sV[k0 + lane / 16][ti * 16 + lane % 16];
                double b = sW[k0 + lane / 16][tj * 16 + lane % 16];
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    double *out = partial + (long long)blockIdx.x * JCI_MAXBASIS * JCI_MAXBASIS;
    for (int t = 0; t < 4; t++) {
        int tile = wave + 4 * t;
        if (tile >= ntile) break;
        int ti = tile / wt, tj = tile % wt;
        for (int v = 0; v < 4; v++) {
            int i = ti * 16 + lane / 16 + 4 * v, j = tj * 16 + lane % 16;
            if (i < m && j < w) out[i * JCI_MAXBASIS + j] = acc[t][v];
        }
    }
}

// one wavefront per output element: lanes take the partials b = lane, lane + 64, ... and a
// butterfly finishes the sum (fixed order, deterministic)
__global__ void k_gram_reduce(int m, int w, int nblocks, const double *__restrict__ partial,
                              double *__restrict__ out)
{
    int idx = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (idx >= m * w) return;
    int i = idx / w, j = idx % w;
    double s = 0.0;
    for (int b = lane; b < nblocks; b += 64)
        s += partial[(long long)b * JCI_MAXBASIS * JCI_MAXBASIS + i * JCI_MAXBASIS + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[i * w + j] = s;
}

// reduction work-groups: 2048 rows each (the chunks are walked sequentially, so shorter
// chunks put more independent loads in flight), at most `cap`
static int reduction_blocks(long long N, int cap)
{
    long long b = (N + 2047) / 2048;
    if (b > cap) b = cap;
    return (int)(b < 1 ? 1 : b);
}

int jk_gram(int m, int w, long long N, const double *V, long long ldv, const double *W,
            long long ldw, double *partial, int nblocks, double *out_dev, hipStream_t st)
{
    if (m < 1 || w < 1 || m > JCI_MAXBASIS || w > JCI_MAXBASIS) {
        jci_set_error("gram: dimensions out of range");
        return -1;
    }
    if ((m + 15) / 16 * ((w + 15) / 16) > 16) {
        jci_set_error("gram: at most 16 output tiles");
        return -1;
    }
    nblocks = reduction_blocks(N, nblocks);
    long long chunk = (N + nblocks - 1) / nblocks;
    chunk = (chunk + GRAM_ROWS - 1) / GRAM_ROWS * GRAM_ROWS;
    int nb = (int)((N + chunk - 1) / chunk);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_gram, dim3(nb), dim3(256), 0, st, m, w, N, V, ldv, W, ldw, partial, chunk);
    hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(m * w, 4)), dim3(256), 0, st, m, w, nb,
                       partial, out_dev);
    LAUNCH_CHECK();
    return 0;
}

// out_j = beta * out_j + sum_i V_i * Y[i*k + j]   (j < k, i < m)
__global__ void k_combine(int m, int k, long long N, const double *__restrict__ V, long long ldv,
                          const double *__restrict__ Y, double *__restrict__ out, long long ldo,
                          double beta)
{
    __shared__ double sY[JCI_MAXBASIS * JCI_MAXK];
    for (int t = threadIdx.x; t < m * k; t += blockDim.x) sY[t] = Y[t];
    __syncthreads();
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double acc[JCI_MAXK];
    for (int j = 0; j < k; j++) acc[j] = beta == 0.0 ? 0.0 : beta * out[(long long)j * ldo + n];
    for (int i = 0; i < m; i++) {
        double vi = V[(long long)i * ldv + n];
        for (int j = 0; j < k; j++) acc[j] += vi * sY[i * k + j];
    }
    for (int j = 0; j < k; j++) out[(long long)j * ldo + n] = acc[j];
}

int jk_combine(int m, int k, long long N, const double *V, long long ldv, const double *Y,
               int ldy, double *out, long long ldo, double beta, hipStream_t st)
{
    (void)ldy;
    if (m < 1 || m > JCI_MAXBASIS || k < 1 || k > JCI_MAXK) {
        jci_set_error("combine: dimensions out of range");
        return -1;
    }
    hipLaunchKernelGGL(k_combine, dim3(grid_for(N, 256)), dim3(256), 0, st, m, k, N, V, ldv, Y,
                       out, ldo, beta);
    LAUNCH_CHECK();
    return 0;
}

// Residuals and preconditioned corrections for k roots:
//   R_r = sum_i Y[i*k+r] (S_i - theta_r V_i),  T_r = -R_r / (hd - theta_r),
// partial[b*k + r] = sum over the block's rows of R_r^2 (fixed-order block reduction).
__global__ __launch_bounds__(256) void k_residual(int m, int k, long long N,
                                                  const double *__restrict__ V,
                                                  const double *__restrict__ S, long long ld,
                                                  const double *__restrict__ Y,
                                                  const double *__restrict__ theta,
                                                  const double *__restrict__ hd,
                                                  double *__restrict__ T,
                                                  double *__restrict__ partial, long long chunk)
{
    __shared__ double sY[JCI_MAXBASIS * JCI_MAXK];
    __shared__ double sT[JCI_MAXK];
    __shared__ double red[JCI_MAXK][256];
    for (int t = threadIdx.x; t < m * k; t += blockDim.x) sY[t] = Y[t];
    if (threadIdx.x < k) sT[threadIdx.x] = theta[threadIdx.x];
    __syncthreads();
    double nrm[JCI_MAXK];
    for (int r = 0; r < k; r++) nrm[r] = 0.0;
    long long r0 = (long long)blockIdx.x * chunk, r1 = r0 + chunk < N ? r0 + chunk : N;
    for (long long n = r0 + threadIdx.x; n < r1; n += blockDim.x) {
        double res[JCI_MAXK];
        for (int r = 0; r < k; r++) res[r] = 0.0;
        // four basis vectors per step: their eight loads are issued together
        int i = 0;
        for (; i + 4 <= m; i += 4) {
            double vi[4], si[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                vi[q] = V[(long long)(i + q) * ld + n];
                si[q] = S[(long long)(i + q) * ld + n];
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                for (int r = 0; r < k; r++) res[r] += sY[(i + q) * k + r] * (si[q] - sT[r] * vi[q]);
        }
        for (; i < m; i++) {
            double vi = V[(long long)i * ld + n], si = S[(long long)i * ld + n];
            for (int r = 0; r < k; r++) res[r] += sY[i * k + r] * (si - sT[r] * vi);
        }
        double h = hd[n];
        for (int r = 0; r < k; r++) {
            nrm[r] += res[r] * res[r];
            double den = h - sT[r];
            if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
            T[(long long)r * ld + n] = -res[r] / den;
        }
    }
    for (int r = 0; r < k; r++) red[r][threadIdx.x] = nrm[r];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int r = 0; r < k; r++) red[r][threadIdx.x] += red[r][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < k) partial[(long long)blockIdx.x * JCI_MAXK + threadIdx.x] = red[threadIdx.x][0];
}

// out[r] = sum_b partial[b*JCI_MAXK + r], one wavefront per r (fixed order)
__global__ void k_partial_reduce(int k, int nblocks, const double *__restrict__ partial,
                                 double *__restrict__ out)
{
    int r = blockIdx.x, lane = threadIdx.x;
    if (r >= k) return;
    double s = 0.0;
    for (int b = lane; b < nblocks; b += 64) s += partial[(long long)b * JCI_MAXK + r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[r] = s;
}

int jk_residual(int m, int k, long long N, const double *V, const double *S, long long ld,
                const double *Y, const double *theta, const double *hd, double *T,
                double *partial, int nblocks, double *norms_dev, hipStream_t st)
{
    if (m < 1 || m > JCI_MAXBASIS || k < 1 || k > JCI_MAXK) {
        jci_set_error("residual: dimensions out of range");
        return -1;
    }
    nblocks = reduction_blocks(N, nblocks);
    long long chunk = (N + nblocks - 1) / nblocks;
    int nb = (int)((N + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, m, k, N, V, S, ld, Y, theta, hd,
                       T, partial, chunk);
    hipLaunchKernelGGL(k_partial_reduce, dim3(k), dim3(64), 0, st, k, nb, partial, norms_dev);
    LAUNCH_CHECK();
    return 0;
}

__global__ __launch_bounds__(256) void k_colnorms(int k, long long N, const double *__restrict__ X,
                                                  long long ld, double *__restrict__ partial,
                                                  long long chunk)
{
    __shared__ double red[JCI_MAXK][256];
    double s[JCI_MAXK];
    for (int r = 0; r < k; r++) s[r] = 0.0;
    long long r0 = (long long)blockIdx.x * chunk, r1 = r0 + chunk < N ? r0 + chunk : N;
    for (long long n = r0 + threadIdx.x; n < r1; n += blockDim.x)
        for (int r = 0; r < k; r++) {
            double v = X[(long long)r * ld + n];
            s[r] += v * v;
        }
    for (int r = 0; r < k; r++) red[r][threadIdx.x] = s[r];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int r = 0; r < k; r++) red[r][threadIdx.x] += red[r][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < k) partial[(long long)blockIdx.x * JCI_MAXK + threadIdx.x] = red[threadIdx.x][0];
}

int jk_colnorms(int k, long long N, const double *X, long long ld, double *partial,
                int nblocks, double *out_dev, hipStream_t st)
{
    if (k < 1 || k > JCI_MAXK) { jci_set_error("colnorms: k out of range"); return -1; }
    nblocks = reduction_blocks(N, nblocks);
    long long chunk = (N + nblocks - 1) / nblocks;
    int nb = (int)((N + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_colnorms, dim3(nb), dim3(256), 0, st, k, N, X, ld, partial, chunk);
    hipLaunchKernelGGL(k_partial_reduce, dim3(k), dim3(64), 0, st, k, nb, partial, out_dev);
    LAUNCH_CHECK();
    return 0;
}

__global__ void k_pack(int k, long long N, const double *__restrict__ src, long long lds,
                       double *__restrict__ dst)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int j = 0; j < k; j++) dst[n * k + j] = src[(long long)j * lds + n];
}

__global__ void k_unpack(int k, long long N, const double *__restrict__ src,
                         double *__restrict__ dst, long long ldd)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int j = 0; j < k; j++) dst[(long long)j * ldd + n] = src[n * k + j];
}

int jk_pack(int k, long long N, const double *src, long long lds, double *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack, dim3(grid_for(N, 256)), dim3(256), 0, st, k, N, src, lds, dst);
    LAUNCH_CHECK();
    return 0;
}

int jk_unpack(int k, long long N, const double *src, double *dst, long long ldd, hipStream_t st)
{
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(N, 256)), dim3(256), 0, st, k, N, src, dst, ldd);
    LAUNCH_CHECK();
    return 0;
}

__global__ void k_scale_cols(int k, long long N, double *__restrict__ X, long long ld,
                             const double *__restrict__ s)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int j = 0; j < k; j++) X[(long long)j * ld + n] *= s[j];
}

int jk_scale_cols(int k, long long N, double *X, long long ld, const double *s, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(N, 256)), dim3(256), 0, st, k, N, X, ld, s);
    LAUNCH_CHECK();
    return 0;
}

// dense H over a small set of determinants (initial-guess block)
__global__ void k_dense_block(int nb, const uint64_t *__restrict__ ma, const uint64_t *__restrict__ mb,
                              const double *__restrict__ m1, const double *__restrict__ m2,
                              double *__restrict__ H)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * nb) return;
    int i = idx / nb, j = idx % nb;
    int dg, ph;
    H[idx] = jci_hij(ma[i], mb[i], ma[j], mb[j], m1, m2, &dg, &ph);
}

int jk_dense_block(int nb, const uint64_t *ma_rows, const uint64_t *mb_rows, int norbs,
                   const double *m1, const double *m2, double *H, hipStream_t st)
{
    (void)norbs;
    hipLaunchKernelGGL(k_dense_block, dim3(grid_for((long long)nb * nb, 256)), dim3(256), 0, st,
                       nb, ma_rows, mb_rows, m1, m2, H);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// Explicit determinant lists (the reference-format drop-in path).

__global__ void k_list_hdiag(long long n, const uint64_t *__restrict__ ma,
                             const uint64_t *__restrict__ mb, const double *__restrict__ m1,
                             const double *__restrict__ m2, double *__restrict__ hd)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hd[i] = jci_diag(ma[i], mb[i], m1, m2);
}

int jk_list_hdiag(long long n, const uint64_t *ma, const uint64_t *mb, int norbs,
                  const double *m1, const double *m2, double *hd, hipStream_t st)
{
    (void)norbs;
    hipLaunchKernelGGL(k_list_hdiag, dim3(grid_for(n, 256)), dim3(256), 0, st, n, ma, mb, m1, m2, hd);
    LAUNCH_CHECK();
    return 0;
}

// Connectivity screen of source/cimapping.c:92 (generate_cimap): for row i, every j >= i
// whose excitation degree popcount((Ia^Ja)&Ia) + popcount((Ib^Jb)&Ib) is <= 2.  One
// thread per row, 256 rows per work-group; candidate determinants are staged in LDS in
// tiles and read as broadcasts, so each tile is fetched once per 256 rows.
#define SCREEN_TILE 512
template <bool FILL>
__global__ __launch_bounds__(256) void k_list_screen(long long n, const uint64_t *__restrict__ ma,
                                                     const uint64_t *__restrict__ mb,
                                                     const long long *__restrict__ rowptr,
                                                     long long *__restrict__ cnt,
                                                     int *__restrict__ col)
{
    __shared__ uint64_t sa[SCREEN_TILE], sb[SCREEN_TILE];
    long long rowmin = (long long)blockIdx.x * 256;
    long long row = rowmin + threadIdx.x;
    bool active = row < n;
    uint64_t ia = active ? ma[row] : 0, ib = active ? mb[row] : 0;
    long long pos = (FILL && active) ? rowptr[row] : 0, count = 0;
    for (long long base = rowmin; base < n; base += SCREEN_TILE) {
        __syncthreads();
        for (int t = threadIdx.x; t < SCREEN_TILE; t += 256) {
            long long c = base + t;
            sa[t] = c < n ? ma[c] : 0;
            sb[t] = c < n ? mb[c] : 0;
        }
        __syncthreads();
        if (!active) continue;
        int lim = (int)((n - base) < SCREEN_TILE ? (n - base) : SCREEN_TILE);
        int t0 = row > base ? (int)(row - base) : 0;
        for (int t = t0; t < lim; t++) {
            int dd = jci_popc((ia ^ sa[t]) & ia) + jci_popc((ib ^ sb[t]) & ib);
            if (dd <= 2) {
                if (FILL) col[pos] = (int)(base + t);
                pos++;
                count++;
            }
        }
    }
    if (!FILL && active) cnt[row] = count;
}

int jk_list_screen_count(long long n, const uint64_t *ma, const uint64_t *mb, long long *cnt,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_list_screen<false>, dim3(grid_for(n, 256)), dim3(256), 0, st, n, ma, mb,
                       (const long long *)nullptr, cnt, (int *)nullptr);
    LAUNCH_CHECK();
    return 0;
}

int jk_list_screen_fill(long long n, const uint64_t *ma, const uint64_t *mb,
                        const long long *rowptr, int *col, hipStream_t st)
{
    hipLaunchKernelGGL(k_list_screen<true>, dim3(grid_for(n, 256)), dim3(256), 0, st, n, ma, mb,
                       rowptr, (long long *)nullptr, col);
    LAUNCH_CHECK();
    return 0;
}

// values of a pair list over an explicit determinant list (row-parallel)
__global__ void k_list_values(long long n, const uint64_t *__restrict__ ma,
                              const uint64_t *__restrict__ mb, const double *__restrict__ m1,
                              const double *__restrict__ m2, const long long *__restrict__ rowptr,
                              const int *__restrict__ col, double *__restrict__ val)
{
    long long i = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (i >= n) return;
    uint64_t ia = ma[i], ib = mb[i];
    for (long long t = rowptr[i] + lane; t < rowptr[i + 1]; t += 64) {
        int j = col[t];
        int dg, ph;
        val[t] = jci_hij(ia, ib, ma[j], mb[j], m1, m2, &dg, &ph);
    }
}

int jk_list_values(long long nnz_rows, long long n, const uint64_t *ma, const uint64_t *mb,
                   int norbs, const double *m1, const double *m2, const long long *rowptr,
                   const int *col, double *val, hipStream_t st)
{
    (void)nnz_rows; (void)norbs;
    hipLaunchKernelGGL(k_list_values, dim3(grid_for(n, 4)), dim3(256), 0, st, n, ma, mb, m1, m2,
                       rowptr, col, val);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// small helpers of the handle API

// masks of determinants idx[t] (or first + t when idx is null), t < count
__global__ void k_det_masks(const SpaceDesc *__restrict__ d, const uint64_t *__restrict__ MA,
                            const uint64_t *__restrict__ MB, const long long *__restrict__ idx,
                            long long first, long long count, uint64_t *__restrict__ oa,
                            uint64_t *__restrict__ ob)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    long long I = idx ? idx[t] : first + t;
    DetRef r = decode(d, I);
    oa[t] = MA[r.ia];
    ob[t] = MB[r.ib];
}

int jk_det_masks(const SpaceDesc *d_desc, const uint64_t *ma, const uint64_t *mb,
                 const long long *idx, long long first, long long count, uint64_t *oa,
                 uint64_t *ob, hipStream_t st)
{
    hipLaunchKernelGGL(k_det_masks, dim3(grid_for(count, 256)), dim3(256), 0, st, d_desc, ma, mb,
                       idx, first, count, oa, ob);
    LAUNCH_CHECK();
    return 0;
}

// V_j <- sum_i V_i Y[i*kc + j] for j < kc <= m, in place: each thread reads its row of all
// m vectors before writing any
__global__ __launch_bounds__(256) void k_rotate_inplace(int m, int kc, long long N,
                                                        double *__restrict__ V, long long ld,
                                                        const double *__restrict__ Y)
{
    __shared__ double sY[JCI_MAXBASIS * JCI_MAXBASIS];
    for (int t = threadIdx.x; t < m * kc; t += blockDim.x) sY[t] = Y[t];
    __syncthreads();
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double v[JCI_MAXBASIS];
#pragma unroll
    for (int i = 0; i < JCI_MAXBASIS; i++)
        if (i < m) v[i] = V[(long long)i * ld + n];
    for (int j = 0; j < kc; j++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < JCI_MAXBASIS; i++)
            if (i < m) acc += v[i] * sY[i * kc + j];
        V[(long long)j * ld + n] = acc;
    }
}

int jk_rotate_inplace(int m, int kc, long long N, double *V, long long ld, const double *Y,
                      hipStream_t st)
{
    if (m < 1 || m > JCI_MAXBASIS || kc < 1 || kc > m) {
        jci_set_error("rotate: dimensions out of range (m=%d kc=%d)", m, kc);
        return -1;
    }
    hipLaunchKernelGGL(k_rotate_inplace, dim3(grid_for(N, 256)), dim3(256), 0, st, m, kc, N, V,
                       ld, Y);
    LAUNCH_CHECK();
    return 0;
}

// V_j[idx[t]] = U[j*nb + t]   (eigenvector j of the guess block)
__global__ void k_scatter_guess(int nb, int nvec, const long long *__restrict__ idx,
                                const double *__restrict__ U, double *__restrict__ V, long long ld)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * nvec) return;
    int j = t / nb, r = t % nb;
    V[(long long)j * ld + idx[r]] = U[(long long)j * nb + r];
}

// The nb lowest diagonal elements of each chunk of JCI_GUESS_CHUNK determinants, ascending
// by value with ties by index (the order of the host selection), one block per chunk:
// bitonic sort of the chunk in LDS (entries past N sort last as +inf).  The host merges the
// chunks' candidates.
__global__ __launch_bounds__(256) void k_lowest_chunks(long long N, const double *__restrict__ hd,
                                                       int nb, double *__restrict__ cv,
                                                       long long *__restrict__ ci)
{
    __shared__ double sv[JCI_GUESS_CHUNK];
    __shared__ long long si[JCI_GUESS_CHUNK];
    const long long base = (long long)blockIdx.x * JCI_GUESS_CHUNK;
    for (int t = threadIdx.x; t < JCI_GUESS_CHUNK; t += 256) {
        const long long g = base + t;
        sv[t] = g < N ? hd[g] : __builtin_huge_val();
        si[t] = g < N ? g : 0x7FFFFFFFFFFFFFFFLL;
    }
    __syncthreads();
    for (int k = 2; k <= JCI_GUESS_CHUNK; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < JCI_GUESS_CHUNK; t += 256) {
                const int q = t ^ j;
                if (q > t) {
                    const bool up = (t & k) == 0;
                    const bool gt = sv[t] > sv[q] || (sv[t] == sv[q] && si[t] > si[q]);
                    if (gt == up) {
                        const double v = sv[t];
                        sv[t] = sv[q];
                        sv[q] = v;
                        const long long i = si[t];
                        si[t] = si[q];
                        si[q] = i;
                    }
                }
            }
            __syncthreads();
        }
    for (int t = threadIdx.x; t < nb; t += 256) {
        cv[(long long)blockIdx.x * nb + t] = sv[t];
        ci[(long long)blockIdx.x * nb + t] = si[t];
    }
}

int jk_lowest_chunks(long long N, const double *hd, int nb, double *cv, long long *ci,
                     hipStream_t st)
{
    if (nb < 1 || nb > JCI_GUESS_CHUNK || N < 1) return -1;
    const long long nch = (N + JCI_GUESS_CHUNK - 1) / JCI_GUESS_CHUNK;
    hipLaunchKernelGGL(k_lowest_chunks, dim3((unsigned)nch), dim3(256), 0, st, N, hd, nb, cv, ci);
    LAUNCH_CHECK();
    return 0;
}

int jk_scatter_guess(int nb, int nvec, const long long *idx, const double *Yb, double *V,
                     long long ld, hipStream_t st)
{
    hipLaunchKernelGGL(k_scatter_guess, dim3(grid_for((long long)nb * nvec, 256)), dim3(256), 0,
                       st, nb, nvec, idx, Yb, V, ld);
    LAUNCH_CHECK();
    return 0;
}

__global__ void k_index_of(const SpaceDesc *__restrict__ d, const unsigned long long *__restrict__ binom,
                           uint64_t a, uint64_t b, long long *out)
{
    int gp, gq;
    long long ia = str_index(d, 0, a, binom, &gp);
    long long ib = str_index(d, 1, b, binom, &gq);
    long long r = -1;
    if (ia >= 0 && ib >= 0 && jci_popc(a) == d->s[0].nelec && jci_popc(b) == d->s[1].nelec &&
        (a & ~d->allmask) == 0 && (b & ~d->allmask) == 0) {
        int bk = d->blkid[gp][gq];
        if (bk >= 0)
            r = d->boff[bk] + (ia - d->s[0].gstart[gp]) * d->s[1].gn[gq] + (ib - d->s[1].gstart[gq]);
    }
    *out = r;
}

int jk_index_of(const SpaceDesc *d_desc, const unsigned long long *binom, uint64_t a,
                uint64_t b, long long *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_index_of, dim3(1), dim3(1), 0, st, d_desc, binom, a, b, out);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------------------------
// measurement helpers

__global__ __launch_bounds__(256) void k_copy(long long n16, const double4 *__restrict__ src,
                                              double4 *__restrict__ dst)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

int jk_copy(long long bytes, const void *src, void *dst, hipStream_t st)
{
    long long n = bytes / 32;
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, st, n, (const double4 *)src, (double4 *)dst);
    LAUNCH_CHECK();
    return 0;
}

__global__ void k_fill_random(long long n, unsigned long long seed, double *__restrict__ x)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed + (unsigned long long)i * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    x[i] = (double)(z >> 11) * 0x1.0p-53 - 0.5;
}

int jk_fill_random(long long n, unsigned long long seed, double *x, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_random, dim3(grid_for(n, 256)), dim3(256), 0, st, n, seed, x);
    LAUNCH_CHECK();
    return 0;
}
