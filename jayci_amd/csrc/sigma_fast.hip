// Matrix-free sigma from per-string excitation tables (gfx950).
//
// Replaces the compute_hv loop of source/action_util.c:121 (c += H v) with the
// string-driven organisation of mpi_source/pdavidson.c:1964-2067: for determinant
// (Ia, Ib) the connected determinants are (Ja, Ib) for alpha singles/doubles of Ia,
// (Ia, Jb) for beta singles/doubles of Ib and (Ja, Jb) for alpha single x beta single.
//
// Excitation tables (one per distinct string set), built once per Hamiltonian:
//   e1[off1[I] .. off1[I+1])  single excitations of string I to strings of the space:
//       {J, info}, info = i | f<<6 | neg<<12 | group(J)<<13
//       v1 = h_if + sum_{k in I, k != i} [(kk|if) - (ki|kf)]   (one-spin part, unsigned)
//   e2[off2[I] .. off2[I+1])  double excitations: {J, info},
//       info = i0 | i1<<6 | f0<<12 | f1<<18 | neg<<24 | group(J)<<25
//       v2 = sign * [(i0 f0|i1 f1) - (i0 f1|i1 f0)]
// The element of a single excitation in determinant context adds the cross-spin term
// sum_{k in other string} (kk|if).
//
// Tiles: one work-group of W threads owns the determinants (Ia, Ib_0 .. Ib_0+W-1) of one
// block; thread t owns (Ia, Ib_0 + t).  Alpha excitations are uniform over the work-group
// (reads of the target row are coalesced); beta excitations are per thread.  For each
// alpha single the integral row (i f | .) is staged in LDS, giving the cross-spin term of
// the alpha single and every alpha-beta double element with one LDS read.
#include "jci_internal.h"
#include "slater.h"
#include "sigma_fast.h"

#define LAUNCH_CHECK2()                                                  \
    do {                                                                 \
        hipError_t e_ = hipGetLastError();                               \
        if (e_ != hipSuccess) {                                          \
            jci_set_error("%s:%d launch: %s", __FILE__, __LINE__,        \
                          hipGetErrorString(e_));                        \
            return -2;                                                   \
        }                                                                \
    } while (0)

static inline unsigned gridn(long long n, int per) { long long g = (n + per - 1) / per; return (unsigned)(g > 0 ? g : 1); }

__device__ __forceinline__ uint64_t shr2(uint64_t x, int s) { return s >= 64 ? 0ULL : (x >> s); }

__device__ __forceinline__ long long str_index2(const SpaceDesc *d, int s, uint64_t m,
                                                const unsigned long long *binom, int *g)
{
    int nd = jci_popc(m & d->dmask), na = jci_popc(m & d->amask), nv = jci_popc(m & d->vmask);
    const SpinDesc &S = d->s[s];
    int gg = -1;
    for (int t = 0; t < S.ngrp; t++)
        if (S.gd[t] == nd && S.ga[t] == na && S.gv[t] == nv) { gg = t; break; }
    if (gg < 0) return -1;
    long long rd = jci_lex_rank(m & d->dmask, d->ndocc, nd, binom);
    long long ra = jci_lex_rank(shr2(m & d->amask, d->ndocc), d->nactv, na, binom);
    long long rv = jci_lex_rank(shr2(m & d->vmask, d->ndocc + d->nactv), d->nvirt, nv, binom);
    *g = gg;
    return S.gstart[gg] + (rd * S.nsa[gg] + ra) * S.nsv[gg] + rv;
}

// ---- excitation tables -------------------------------------------------------------------

template <bool FILL>
__global__ void k_exc(const SpaceDesc *__restrict__ d, int spin, long long nstr,
                      const uint64_t *__restrict__ M, const unsigned long long *__restrict__ binom,
                      const double *__restrict__ m1, const double *__restrict__ m2,
                      long long *__restrict__ cnt1, long long *__restrict__ cnt2,
                      const long long *__restrict__ off1, const long long *__restrict__ off2,
                      int2 *__restrict__ e1, double *__restrict__ v1, int2 *__restrict__ e2,
                      double *__restrict__ v2)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nstr) return;
    const uint64_t m = M[I], vir = d->allmask & ~m;
    long long n1 = 0, n2 = 0, p1 = FILL ? off1[I] : 0, p2 = FILL ? off2[I] : 0;
    for (uint64_t x = m; x; x &= x - 1) {
        int i = jci_ctz(x);
        for (uint64_t y = vir; y; y &= y - 1) {
            int f = jci_ctz(y);
            int g;
            long long J = str_index2(d, spin, m ^ (1ULL << i) ^ (1ULL << f), binom, &g);
            if (J < 0) continue;
            if (FILL) {
                int neg = jci_sign(m, i, f) < 0;
                e1[p1] = make_int2((int)J, i | (f << 6) | (neg << 12) | (g << 13));
                double v = jci_h1(m1, i, f);
                for (uint64_t z = m & ~(1ULL << i); z; z &= z - 1) {
                    int k = jci_ctz(z);
                    v += jci_eri(m2, k, k, i, f) - jci_eri(m2, k, i, k, f);
                }
                v1[p1] = v;
                p1++;
            }
            n1++;
        }
    }
    for (uint64_t x0 = m; x0; x0 &= x0 - 1) {
        int i0 = jci_ctz(x0);
        for (uint64_t x1 = x0 & (x0 - 1); x1; x1 &= x1 - 1) {
            int i1 = jci_ctz(x1);
            for (uint64_t y0 = vir; y0; y0 &= y0 - 1) {
                int f0 = jci_ctz(y0);
                for (uint64_t y1 = y0 & (y0 - 1); y1; y1 &= y1 - 1) {
                    int f1 = jci_ctz(y1);
                    int g;
                    long long J = str_index2(d, spin, m ^ (1ULL << i0) ^ (1ULL << i1) ^
                                                          (1ULL << f0) ^ (1ULL << f1),
                                             binom, &g);
                    if (J < 0) continue;
                    if (FILL) {
                        int ph;
                        double v = jci_double_same(m, i0, i1, f0, f1, m2, &ph);
                        e2[p2] = make_int2((int)J, i0 | (i1 << 6) | (f0 << 12) | (f1 << 18) |
                                                       ((ph < 0) << 24) | (g << 25));
                        v2[p2] = v;
                        p2++;
                    }
                    n2++;
                }
            }
        }
    }
    if (!FILL) { cnt1[I] = n1; cnt2[I] = n2; }
}

int jk_exc_count(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                 const unsigned long long *binom, long long *cnt1, long long *cnt2, hipStream_t st)
{
    hipLaunchKernelGGL(k_exc<false>, dim3(gridn(nstr, 64)), dim3(64), 0, st, d, spin, nstr, M,
                       binom, (const double *)nullptr, (const double *)nullptr, cnt1, cnt2,
                       (const long long *)nullptr, (const long long *)nullptr, (int2 *)nullptr,
                       (double *)nullptr, (int2 *)nullptr, (double *)nullptr);
    LAUNCH_CHECK2();
    return 0;
}

int jk_exc_fill(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                const unsigned long long *binom, const double *m1, const double *m2,
                const long long *off1, const long long *off2, int2 *e1, double *v1, int2 *e2,
                double *v2, hipStream_t st)
{
    hipLaunchKernelGGL(k_exc<true>, dim3(gridn(nstr, 64)), dim3(64), 0, st, d, spin, nstr, M,
                       binom, m1, m2, (long long *)nullptr, (long long *)nullptr, off1, off2,
                       e1, v1, e2, v2);
    LAUNCH_CHECK2();
    return 0;
}

// square copy of the packed two-electron integrals: sq[P*m1 + Q] = (P|Q), pair indices
// P = tri(i,j), so row P is contiguous
__global__ void k_m2_square(int m1, const double *__restrict__ m2, double *__restrict__ sq)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)m1 * m1) return;
    int P = (int)(t / m1), Q = (int)(t % m1);
    sq[t] = m2[jci_tri(P, Q)];
}

int jk_m2_square(int m1, const double *m2, double *sq, hipStream_t st)
{
    hipLaunchKernelGGL(k_m2_square, dim3(gridn((long long)m1 * m1, 256)), dim3(256), 0, st, m1,
                       m2, sq);
    LAUNCH_CHECK2();
    return 0;
}

// ---- sigma tile kernel ------------------------------------------------------------------

__device__ __forceinline__ int tile_block(const long long *__restrict__ toff, int nblk, long long t)
{
    int lo = 0, hi = nblk - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (toff[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename AT, typename XT, int K, int W>
__global__ __launch_bounds__(W) void k_sigma_fast(FastArgs a, const XT *__restrict__ x,
                                                  double *__restrict__ y)
{
    extern __shared__ unsigned char smem_raw[];
    AT *g = (AT *)smem_raw;              // staged integral row (i f | .), m1 entries
    AT *Fa = g + a.m1;                   // sum_{k in Ia} (kk | .), m1 entries
    const SpaceDesc *__restrict__ d = a.d;
    const int tid = threadIdx.x;
    const long long tile = (long long)blockIdx.x;
    const int b = tile_block(a.tile_off, d->nblk, tile);
    const int p = d->bp[b], q = d->bq[b];
    const long long nq = d->s[1].gn[q];
    const long long nch = (nq + W - 1) / W;
    const long long rel = tile - a.tile_off[b];
    const long long ia_loc = rel / nch, chunk = rel - ia_loc * nch;
    const long long ib_loc = chunk * W + tid;
    const bool active = ib_loc < nq;
    const long long Ia = d->s[0].gstart[p] + ia_loc;
    const long long Ib = d->s[1].gstart[q] + (active ? ib_loc : 0);
    const uint64_t ma = a.ma[Ia];
    const uint64_t mb = a.mb[Ib];
    const long long I = d->boff[b] + ia_loc * nq + ib_loc;

    // Fa[P] = sum over alpha electrons k of (kk|P)
    for (int P = tid; P < a.m1; P += W) {
        double s = 0.0;
        for (uint64_t z = ma; z; z &= z - 1) {
            int k = jci_ctz(z);
            s += a.m2sq[(long long)jci_tri(k, k) * a.m1 + P];
        }
        Fa[P] = (AT)s;
    }
    __syncthreads();

    AT acc[K];
    if (active) {
        AT h = (AT)a.hd[I];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = h * (AT)x[I * K + k];
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = (AT)0;
    }
    const SpinDesc &SA = d->s[0], &SB = d->s[1];
    const ExcView &TA = a.ta, &TB = a.tb;

    // beta excitations of Ib (per thread)
    if (active) {
        for (long long e = TB.off1[Ib]; e < TB.off1[Ib + 1]; e++) {
            int2 en = TB.e1[e];
            int gq = (en.y >> 13) & 31;
            int bk = d->blkid[p][gq];
            if (bk < 0) continue;
            int i = en.y & 63, f = (en.y >> 6) & 63;
            long long J = d->boff[bk] + ia_loc * SB.gn[gq] + (en.x - SB.gstart[gq]);
            AT v = (AT)TB.v1[e] + Fa[jci_tri(i, f)];
            if (en.y & (1 << 12)) v = -v;
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
        }
        for (long long e = TB.off2[Ib]; e < TB.off2[Ib + 1]; e++) {
            int2 en = TB.e2[e];
            int gq = (en.y >> 25) & 31;
            int bk = d->blkid[p][gq];
            if (bk < 0) continue;
            long long J = d->boff[bk] + ia_loc * SB.gn[gq] + (en.x - SB.gstart[gq]);
            AT v = (AT)TB.v2[e];
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
        }
    }

    // alpha doubles (uniform over the work-group)
    for (long long e = TA.off2[Ia]; e < TA.off2[Ia + 1]; e++) {
        int2 en = TA.e2[e];
        int gp = (en.y >> 25) & 31;
        int bk = d->blkid[gp][q];
        if (bk < 0 || !active) continue;
        long long J = d->boff[bk] + (en.x - SA.gstart[gp]) * nq + ib_loc;
        AT v = (AT)TA.v2[e];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
    }

    // alpha singles (uniform): own element and products with beta singles
    const long long b1 = TB.off1[Ib], e1end = TB.off1[Ib + 1];
    for (long long e = TA.off1[Ia]; e < TA.off1[Ia + 1]; e++) {
        int2 en = TA.e1[e];
        int i = en.y & 63, f = (en.y >> 6) & 63, gp = (en.y >> 13) & 31;
        bool neg = en.y & (1 << 12);
        long long P = jci_tri(i, f);
        __syncthreads();
        for (int Q = tid; Q < a.m1; Q += W) g[Q] = (AT)a.m2sq[P * a.m1 + Q];
        __syncthreads();
        if (!active) continue;
        long long ja_loc = en.x - SA.gstart[gp];
        int bk = d->blkid[gp][q];
        if (bk >= 0) {
            AT cross = (AT)0;
            for (uint64_t z = mb; z; z &= z - 1) {
                int k = jci_ctz(z);
                cross += g[jci_tri(k, k)];
            }
            AT v = (AT)TA.v1[e] + cross;
            if (neg) v = -v;
            long long J = d->boff[bk] + ja_loc * nq + ib_loc;
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
        }
        for (long long e2 = b1; e2 < e1end; e2++) {
            int2 eb = TB.e1[e2];
            int gq = (eb.y >> 13) & 31;
            int bk2 = d->blkid[gp][gq];
            if (bk2 < 0) continue;
            int kk = eb.y & 63, ll = (eb.y >> 6) & 63;
            AT v = g[jci_tri(kk, ll)];
            if (neg != (bool)(eb.y & (1 << 12))) v = -v;
            long long J = d->boff[bk2] + ja_loc * SB.gn[gq] + (eb.x - SB.gstart[gq]);
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
        }
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < K; k++) y[I * K + k] = (double)acc[k];
    }
}

template <typename AT, typename XT, int W>
static int launch_fast(const FastArgs &a, int nvec, const XT *x, double *y, hipStream_t st)
{
    size_t shm = sizeof(AT) * 2 * (size_t)a.m1;
    dim3 g((unsigned)a.ntiles), b(W);
    switch (nvec) {
    case 1: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 1, W>), g, b, shm, st, a, x, y); break;
    case 2: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 2, W>), g, b, shm, st, a, x, y); break;
    case 3: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 3, W>), g, b, shm, st, a, x, y); break;
    case 4: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 4, W>), g, b, shm, st, a, x, y); break;
    case 5: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 5, W>), g, b, shm, st, a, x, y); break;
    case 6: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 6, W>), g, b, shm, st, a, x, y); break;
    case 7: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 7, W>), g, b, shm, st, a, x, y); break;
    case 8: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 8, W>), g, b, shm, st, a, x, y); break;
    default: jci_set_error("sigma_fast: nvec must be 1..8"); return -1;
    }
    LAUNCH_CHECK2();
    return 0;
}

int jk_sigma_fast(int prec, int nvec, const FastArgs &a, const void *x, double *y, hipStream_t st)
{
    if (a.ntiles <= 0) return 0;
    if (a.W == 64) {
        if (prec == 64) return launch_fast<double, double, 64>(a, nvec, (const double *)x, y, st);
        return launch_fast<float, float, 64>(a, nvec, (const float *)x, y, st);
    }
    if (a.W == 128) {
        if (prec == 64) return launch_fast<double, double, 128>(a, nvec, (const double *)x, y, st);
        return launch_fast<float, float, 128>(a, nvec, (const float *)x, y, st);
    }
    if (prec == 64) return launch_fast<double, double, 256>(a, nvec, (const double *)x, y, st);
    return launch_fast<float, float, 256>(a, nvec, (const float *)x, y, st);
}
