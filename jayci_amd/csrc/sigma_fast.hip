// Matrix-free sigma from per-string excitation tables (gfx950).
//
// Replaces the compute_hv loop of source/action_util.c:121 (c += H v) with the
// string-driven organisation of mpi_source/pdavidson.c:1964-2067: for determinant
// (Ia, Ib) the connected determinants are (Ja, Ib) for alpha singles/doubles of Ia,
// (Ia, Jb) for beta singles/doubles of Ib and (Ja, Jb) for alpha single x beta single.
//
// Excitation tables (one per distinct string set), built once per Hamiltonian:
//   e1[off1[I] .. off1[I+1])  single excitations of string I to strings of the space:
//       {J, info}, info = i | f<<6 | neg<<12 | group(J)<<13 | tri(i,f)<<18
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

// Counting pass: per-string totals and per-target-group offsets (goff rows of G+1);
// fill pass: entries written grouped by target group, in (i, f) enumeration order inside
// a group.
template <bool FILL>
__global__ void k_exc(const SpaceDesc *__restrict__ d, int spin, long long nstr,
                      const uint64_t *__restrict__ M, const unsigned long long *__restrict__ binom,
                      const double *__restrict__ m1, const double *__restrict__ m2,
                      long long *__restrict__ cnt1, long long *__restrict__ cnt2,
                      int *__restrict__ goff1, int *__restrict__ goff2,
                      const long long *__restrict__ off1, const long long *__restrict__ off2,
                      int2 *__restrict__ e1, double *__restrict__ v1, int2 *__restrict__ e2,
                      double *__restrict__ v2)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nstr) return;
    const int G = d->s[spin].ngrp;
    const uint64_t m = M[I], vir = d->allmask & ~m;
    int c1[JCI_MAXG + 1], c2[JCI_MAXG + 1];
    for (int g = 0; g <= G; g++) { c1[g] = 0; c2[g] = 0; }
    if (FILL) {
        for (int g = 0; g <= G; g++) {
            c1[g] = goff1[I * (G + 1) + g];
            c2[g] = goff2[I * (G + 1) + g];
        }
    }
    const long long b1 = FILL ? off1[I] : 0, b2 = FILL ? off2[I] : 0;
    for (uint64_t x = m; x; x &= x - 1) {
        int i = jci_ctz(x);
        for (uint64_t y = vir; y; y &= y - 1) {
            int f = jci_ctz(y);
            int g;
            long long J = str_index2(d, spin, m ^ (1ULL << i) ^ (1ULL << f), binom, &g);
            if (J < 0) continue;
            if (FILL) {
                long long p1 = b1 + c1[g];
                int neg = jci_sign(m, i, f) < 0;
                e1[p1] = make_int2((int)J, i | (f << 6) | (neg << 12) | (g << 13) |
                                               ((int)jci_tri(i, f) << 18));
                double v = jci_h1(m1, i, f);
                for (uint64_t z = m & ~(1ULL << i); z; z &= z - 1) {
                    int k = jci_ctz(z);
                    v += jci_eri(m2, k, k, i, f) - jci_eri(m2, k, i, k, f);
                }
                v1[p1] = v;
            }
            c1[g]++;
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
                        long long p2 = b2 + c2[g];
                        int ph;
                        double v = jci_double_same(m, i0, i1, f0, f1, m2, &ph);
                        e2[p2] = make_int2((int)J, i0 | (i1 << 6) | (f0 << 12) | (f1 << 18) |
                                                       ((ph < 0) << 24) | (g << 25));
                        v2[p2] = v;
                    }
                    c2[g]++;
                }
            }
        }
    }
    if (!FILL) {
        long long t1 = 0, t2 = 0;
        for (int g = 0; g < G; g++) {
            goff1[I * (G + 1) + g] = (int)t1;
            goff2[I * (G + 1) + g] = (int)t2;
            t1 += c1[g];
            t2 += c2[g];
        }
        goff1[I * (G + 1) + G] = (int)t1;
        goff2[I * (G + 1) + G] = (int)t2;
        cnt1[I] = t1;
        cnt2[I] = t2;
    }
}

int jk_exc_count(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                 const unsigned long long *binom, long long *cnt1, long long *cnt2,
                 int *goff1, int *goff2, hipStream_t st)
{
    hipLaunchKernelGGL(k_exc<false>, dim3(gridn(nstr, 64)), dim3(64), 0, st, d, spin, nstr, M,
                       binom, (const double *)nullptr, (const double *)nullptr, cnt1, cnt2,
                       goff1, goff2, (const long long *)nullptr, (const long long *)nullptr,
                       (int2 *)nullptr, (double *)nullptr, (int2 *)nullptr, (double *)nullptr);
    LAUNCH_CHECK2();
    return 0;
}

int jk_exc_fill(const SpaceDesc *d, int spin, long long nstr, const uint64_t *M,
                const unsigned long long *binom, const double *m1, const double *m2,
                const long long *off1, const long long *off2, const int *goff1,
                const int *goff2, int2 *e1, double *v1, int2 *e2, double *v2, hipStream_t st)
{
    hipLaunchKernelGGL(k_exc<true>, dim3(gridn(nstr, 64)), dim3(64), 0, st, d, spin, nstr, M,
                       binom, m1, m2, (long long *)nullptr, (long long *)nullptr,
                       (int *)goff1, (int *)goff2, off1, off2, e1, v1, e2, v2);
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

// space lookups cached in LDS by every work-group
struct SpaceLds {
    int blkid[JCI_MAXG][JCI_MAXG];
    long long boff[JCI_MAXB + 1];
    long long gsA[JCI_MAXG + 1], gnA[JCI_MAXG], gsB[JCI_MAXG + 1], gnB[JCI_MAXG];
    int GA, GB;
};

__device__ __forceinline__ void load_space_lds(const SpaceDesc *__restrict__ d, SpaceLds *L, int tid, int nt)
{
    for (int t = tid; t < JCI_MAXG * JCI_MAXG; t += nt) L->blkid[t / JCI_MAXG][t % JCI_MAXG] = d->blkid[t / JCI_MAXG][t % JCI_MAXG];
    for (int t = tid; t <= d->nblk; t += nt) L->boff[t] = d->boff[t];
    for (int t = tid; t < JCI_MAXG; t += nt) {
        L->gsA[t] = d->s[0].gstart[t]; L->gnA[t] = d->s[0].gn[t];
        L->gsB[t] = d->s[1].gstart[t]; L->gnB[t] = d->s[1].gn[t];
    }
    if (tid == 0) { L->GA = d->s[0].ngrp; L->GB = d->s[1].ngrp; }
}

// Alpha singles are taken A at a time: the A integral rows (P_a | .) are staged in LDS
// together, and each beta-list entry a lane loads serves all A of them (A independent
// gathers of x per entry instead of one entry load per gather).  Singles whose target
// group pairs with no beta group reachable from q are skipped without staging.
// XCD-aware tile order: work-groups are dealt to the 8 XCDs round-robin by blockIdx; with
// G > 1, runs of G consecutive tiles (the row segments of one alpha string and its
// neighbours, which read the same vector rows) go to one XCD and share its L2, and the runs
// are dealt round-robin so that every XCD still gets a share of every block.  The grid is
// padded to a multiple of 8 G; tiles past the end return at once.
__device__ __forceinline__ long long xcd_tile(long long b, int G)
{
    if (G <= 1) return b;
    const long long slot = b >> 3, x = b & 7;
    return (slot / G) * (8LL * G) + x * G + slot % G;
}

static inline unsigned xcd_grid(long long ntiles, int G)
{
    if (G <= 1) return (unsigned)ntiles;
    const long long q = 8LL * G;
    return (unsigned)((ntiles + q - 1) / q * q);
}

// acc += sum over table entries [e0, e1) (uniform over the work-group) of v[e] times the
// vector at row e.x of the target group: x[(base + e.x * nq) * K + k].  D entries per round
// (D scalar entry loads, then D coalesced row loads in flight) keep the loads' latencies
// overlapped; the tail runs one at a time.
template <typename AT, typename XT, int K>
__device__ __forceinline__ void uniform_rows(const int2 *__restrict__ e, const double *__restrict__ v,
                                             long long e0, long long e1, long long base,
                                             long long nq, const XT *__restrict__ x, AT *acc)
{
    constexpr int D = 16 / K >= 8 ? 8 : (16 / K >= 4 ? 4 : 2);
    long long i = e0;
    for (; i + D <= e1; i += D) {
        long long J[D];
        AT c[D];
#pragma unroll
        for (int u = 0; u < D; u++) {
            J[u] = base + (long long)e[i + u].x * nq;
            c[u] = (AT)v[i + u];
        }
        XT xv[D][K];
#pragma unroll
        for (int u = 0; u < D; u++)
#pragma unroll
            for (int k = 0; k < K; k++) xv[u][k] = x[J[u] * K + k];
#pragma unroll
        for (int u = 0; u < D; u++)
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += c[u] * (AT)xv[u][k];
    }
    for (; i < e1; i++) {
        const long long J = base + (long long)e[i].x * nq;
        const AT c = (AT)v[i];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += c * (AT)x[J * K + k];
    }
}

// position of tile `rel` of a block of nrow uniform strings x nch lane chunks: row-major
// (the chunks of one string consecutive) or, with cm, chunk-major (consecutive strings on
// one chunk: neighbouring strings share most excitation targets, so consecutive tiles read
// the same vector rows over the same columns)
__device__ __forceinline__ void tile_pos(long long rel, long long nch, long long nrow, int cm,
                                         long long *row, long long *chunk)
{
    if (cm) {
        *chunk = rel / nrow;
        *row = rel - *chunk * nrow;
    } else {
        *row = rel / nch;
        *chunk = rel - *row * nch;
    }
}

__device__ __forceinline__ int tri32(int a, int b)
{
    return a >= b ? ((a * (a + 1)) >> 1) + b : ((b * (b + 1)) >> 1) + a;
}

template <typename AT, typename XT, int K, int W, int A, bool BSIDE>
__global__ __launch_bounds__(W) void k_sigma_fast(FastArgs a, const XT *__restrict__ x,
                                                  double *__restrict__ y)
{
    extern __shared__ unsigned char smem_raw[];
    __shared__ SpaceLds L;
    AT *g = (AT *)smem_raw;              // A staged integral rows (P_a | .), m1 entries each;
                                         // before the alpha singles: Fa
    const SpaceDesc *__restrict__ d = a.d;
    const int tid = threadIdx.x;
    const int m1 = a.m1;
    const long long tile = xcd_tile(blockIdx.x, a.xg);
    if (tile >= a.ntiles) return;
    const int b = tile_block(a.tile_off, d->nblk, tile);
    const int p = d->bp[b], q = d->bq[b];
    const long long nq = d->s[1].gn[q];
    const long long nch = (nq + W - 1) / W;
    const long long rel = tile - a.tile_off[b];
    long long ia_loc, chunk;
    tile_pos(rel, nch, d->s[0].gn[p], a.cm, &ia_loc, &chunk);
    const long long ib_loc = chunk * W + tid;
    const bool active = ib_loc < nq;
    const long long Ia = d->s[0].gstart[p] + ia_loc;
    const long long Ib = d->s[1].gstart[q] + (active ? ib_loc : 0);
    const uint64_t ma = a.ma[Ia];
    const uint64_t mb = a.mb[Ib];
    const long long I = d->boff[b] + ia_loc * nq + ib_loc;
    load_space_lds(d, &L, tid, W);

    // Fa[P] = sum over alpha electrons k of (kk|P)
    if (BSIDE) {
        for (int P = tid; P < m1; P += W) {
            double s = 0.0;
            for (uint64_t z = ma; z; z &= z - 1) {
                int k = jci_ctz(z);
                s += a.m2sq[(long long)tri32(k, k) * m1 + P];
            }
            g[P] = (AT)s;
        }
    }
    __syncthreads();

    // output position (and the diagonal's): I itself, or the same determinant in the
    // transposed layout when this descriptor is the output's transpose
    const long long O = a.tout ? d->boff[b] + ib_loc * d->s[0].gn[p] + ia_loc : I;
    AT acc[K];
    if (active) {
        AT h = (AT)a.hd[O];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = h * (AT)x[I * K + k];
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = (AT)0;
    }
    const ExcView &TA = a.ta, &TB = a.tb;
    const int GA = L.GA, GB = L.GB;

    // beta excitations of Ib (per thread), only towards groups paired with p (without
    // BSIDE the transposed pass k_sigma_fast_beta adds them)
    if (BSIDE && active) {
        const long long o1 = TB.off1[Ib], o2 = TB.off2[Ib];
        const int *go1 = TB.goff1 + Ib * (GB + 1), *go2 = TB.goff2 + Ib * (GB + 1);
        for (int gq = 0; gq < GB; gq++) {
            int bk = L.blkid[p][gq];
            if (bk < 0) continue;
            const long long base = L.boff[bk] + ia_loc * L.gnB[gq] - L.gsB[gq];
            for (long long e = o1 + go1[gq]; e < o1 + go1[gq + 1]; e++) {
                int2 en = TB.e1[e];
                AT v = (AT)TB.v1[e] + g[(en.y >> 18) & 4095];
                if (en.y & (1 << 12)) v = -v;
                long long J = base + en.x;
#pragma unroll
                for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
            }
            for (long long e = o2 + go2[gq]; e < o2 + go2[gq + 1]; e++) {
                AT v = (AT)TB.v2[e];
                long long J = base + TB.e2[e].x;
#pragma unroll
                for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
            }
        }
    }

    // alpha doubles (uniform over the work-group), only towards groups paired with q
    {
        const long long o2 = TA.off2[Ia];
        const int *go2 = TA.goff2 + Ia * (GA + 1);
        for (int gp = 0; gp < GA; gp++) {
            int bk = L.blkid[gp][q];
            if (bk < 0 || !active) continue;
            const long long base = L.boff[bk] - L.gsA[gp] * nq + ib_loc;
            uniform_rows<AT, XT, K>(TA.e2, TA.v2, o2 + go2[gp], o2 + go2[gp + 1], base, nq, x, acc);
        }
    }

    // alpha singles, A at a time within each target group: own element (with the
    // cross-spin term over the beta electrons) and products with the beta singles
    const long long oa1 = TA.off1[Ia];
    const int *goa1 = TA.goff1 + Ia * (GA + 1);
    const long long ob1 = TB.off1[Ib];
    const int *gob1 = TB.goff1 + Ib * (GB + 1);
    for (int gp = 0; gp < GA; gp++) {
        const int bown = L.blkid[gp][q];
        bool any = bown >= 0;
        // pairs the row's beta side reaches: internal ones only while neither the row's beta
        // group nor a beta group paired with gp holds external electrons (the staged rows
        // then stop at the internal pairs, which lead (P|Q))
        int qlim = d->s[1].gv[q] == 0 ? a.m1int : m1;
        for (int gq = 0; gq < GB; gq++) {
            if (L.blkid[gp][gq] < 0) continue;
            any = true;
            if (d->s[1].gv[gq] > 0) qlim = m1;
        }
        if (!any) continue;
        const long long s1 = oa1 + goa1[gp + 1];
        for (long long c0 = oa1 + goa1[gp]; c0 < s1; c0 += A) {
            const int na = (int)(s1 - c0 < A ? s1 - c0 : A);
            long long jl[A];
            int ng[A];
            AT v1[A];
#pragma unroll
            for (int t = 0; t < A; t++) {
                const int2 en = TA.e1[c0 + (t < na ? t : 0)];
                jl[t] = en.x - L.gsA[gp];
                ng[t] = (en.y >> 12) & 1;
                v1[t] = (AT)TA.v1[c0 + (t < na ? t : 0)];
            }
            __syncthreads();
            {
                // A rows, four elements per thread and row in flight per round
                long long prow[A];
#pragma unroll
                for (int t = 0; t < A; t++)
                    prow[t] = (long long)((TA.e1[c0 + (t < na ? t : 0)].y >> 18) & 4095) * m1;
                for (int Q0 = tid; Q0 < qlim; Q0 += 4 * W) {
                    double v[A][4];
#pragma unroll
                    for (int t = 0; t < A; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int Q = Q0 + r * W;
                            v[t][r] = (t < na && Q < qlim) ? a.m2sq[prow[t] + Q] : 0.0;
                        }
#pragma unroll
                    for (int t = 0; t < A; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int Q = Q0 + r * W;
                            if (t < na && Q < qlim) g[t * m1 + Q] = (AT)v[t][r];
                        }
                }
            }
            __syncthreads();
            if (!active) continue;
            if (bown >= 0) {
#pragma unroll
                for (int t = 0; t < A; t++) {
                    if (t >= na) break;
                    AT cross = (AT)0;
                    for (uint64_t z = mb; z; z &= z - 1) {
                        int k = jci_ctz(z);
                        cross += g[t * m1 + tri32(k, k)];
                    }
                    AT v = v1[t] + cross;
                    if (ng[t]) v = -v;
                    const long long J = L.boff[bown] + jl[t] * nq + ib_loc;
#pragma unroll
                    for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
                }
            }
            // the chunk's staged rows beyond na hold stale (possibly never written) LDS:
            // their coefficients are selected to zero, and their gathers repeat row 0's
            // (in bounds)
            for (int gq = 0; gq < GB; gq++) {
                const int bk2 = L.blkid[gp][gq];
                if (bk2 < 0) continue;
                long long base[A];
#pragma unroll
                for (int t = 0; t < A; t++)
                    base[t] = L.boff[bk2] + jl[t < na ? t : 0] * L.gnB[gq] - L.gsB[gq];
                // U beta singles per round: U * A gathers in flight, and behind them the next
                // round's U entry loads
                constexpr int U = 24 / (A * K) >= 4 ? 4 : (24 / (A * K) >= 2 ? 2 : 1);
                const long long e_end = ob1 + gob1[gq + 1];
                long long e2 = ob1 + gob1[gq];
                if (e2 + U <= e_end) {
                    int2 ebn[U];
#pragma unroll
                    for (int u = 0; u < U; u++) ebn[u] = TB.e1[e2 + u];
                    for (; e2 + U <= e_end; e2 += U) {
                        int2 eb[U];
#pragma unroll
                        for (int u = 0; u < U; u++) eb[u] = ebn[u];
                        XT xv[U][A][K];
#pragma unroll
                        for (int u = 0; u < U; u++)
#pragma unroll
                            for (int t = 0; t < A; t++) {
                                const XT *xp = x + (base[t] + eb[u].x) * K;
#pragma unroll
                                for (int k = 0; k < K; k++) xv[u][t][k] = xp[k];
                            }
                        if (e2 + 2 * U <= e_end) {
#pragma unroll
                            for (int u = 0; u < U; u++) ebn[u] = TB.e1[e2 + U + u];
                        }
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int Q = (eb[u].y >> 18) & 4095;
                            const int nb = (eb[u].y >> 12) & 1;
#pragma unroll
                            for (int t = 0; t < A; t++) {
                                AT v = t < na ? g[t * m1 + Q] : (AT)0;
                                if (ng[t] != nb) v = -v;
#pragma unroll
                                for (int k = 0; k < K; k++) acc[k] += v * (AT)xv[u][t][k];
                            }
                        }
                    }
                }
                for (; e2 < e_end; e2++) {
                    const int2 eb = TB.e1[e2];
                    const int Q = (eb.y >> 18) & 4095;
                    const int nb = (eb.y >> 12) & 1;
                    XT xv[A][K];
#pragma unroll
                    for (int t = 0; t < A; t++) {
                        const XT *xp = x + (base[t] + eb.x) * K;
#pragma unroll
                        for (int k = 0; k < K; k++) xv[t][k] = xp[k];
                    }
#pragma unroll
                    for (int t = 0; t < A; t++) {
                        AT v = t < na ? g[t * m1 + Q] : (AT)0;
                        if (ng[t] != nb) v = -v;
#pragma unroll
                        for (int k = 0; k < K; k++) acc[k] += v * (AT)xv[t][k];
                    }
                }
            }
        }
    }
    if (active) {
        if (a.ldy) {
#pragma unroll
            for (int k = 0; k < K; k++) y[k * a.ldy + O] = (double)acc[k];
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) y[O * K + k] = (double)acc[k];
        }
    }
}

// alpha singles per staged chunk: 4 while four integral rows fit 64 KB of LDS, else 2
static inline int fast_chunk(int m1, size_t esz) { return 4 * (size_t)m1 * esz <= 65536 ? 4 : 2; }

// ---- the beta side as a pass over the transposed vector ------------------------------------
// With the space transposed (spins swapped: blocks stored beta-major, see
// jci_transpose_desc), the beta singles' own elements and the beta doubles of the original
// space are the alpha-side work of the transposed one: uniform over a work-group (its
// beta string), with coalesced reads of the transposed vector.  The result is added to the
// original-layout sigma.  Cross-spin term of a single i -> f: sum over the other spin's
// electrons k of (if|kk), read from row tri(i,f) of the square integrals.
template <typename AT, typename XT, int K, int W>
__global__ __launch_bounds__(W) void k_sigma_fast_beta(FastArgs a, const XT *__restrict__ xT,
                                                       double *__restrict__ y)
{
    __shared__ SpaceLds L;
    const SpaceDesc *__restrict__ d = a.d;          // transposed descriptor
    const int tid = threadIdx.x;
    const int m1 = a.m1;
    const long long tile = xcd_tile(blockIdx.x, a.xg);
    if (tile >= a.ntiles) return;
    const int b = tile_block(a.tile_off, d->nblk, tile);
    const int p = d->bp[b], q = d->bq[b];           // beta group, alpha group
    const long long nq = d->s[1].gn[q];             // alpha strings of the block
    const long long nch = (nq + W - 1) / W;
    const long long rel = tile - a.tile_off[b];
    long long ia_loc, chunk;                                          // beta string
    tile_pos(rel, nch, d->s[0].gn[p], a.cm, &ia_loc, &chunk);
    const long long ib_loc = chunk * W + tid;                         // alpha string
    const bool active = ib_loc < nq;
    const long long Ia = d->s[0].gstart[p] + ia_loc;
    const long long Ib = d->s[1].gstart[q] + (active ? ib_loc : 0);
    const uint64_t mb = a.mb[Ib];
    load_space_lds(d, &L, tid, W);
    __syncthreads();
    if (!active) return;
    AT acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = (AT)0;
    const ExcView &TA = a.ta;
    const int GA = L.GA;
    // doubles of the (beta) string, uniform over the work-group
    {
        const long long o2 = TA.off2[Ia];
        const int *go2 = TA.goff2 + Ia * (GA + 1);
        for (int gp = 0; gp < GA; gp++) {
            int bk = L.blkid[gp][q];
            if (bk < 0) continue;
            const long long base = L.boff[bk] - L.gsA[gp] * nq + ib_loc;
            uniform_rows<AT, XT, K>(TA.e2, TA.v2, o2 + go2[gp], o2 + go2[gp + 1], base, nq, xT, acc);
        }
    }
    // singles of the (beta) string: own element with the cross-spin term (its loads issued
    // eight at a time, in flight together)
    {
        const long long o1 = TA.off1[Ia];
        const int *go1 = TA.goff1 + Ia * (GA + 1);
        for (int gp = 0; gp < GA; gp++) {
            int bk = L.blkid[gp][q];
            if (bk < 0) continue;
            const long long base = L.boff[bk] - L.gsA[gp] * nq + ib_loc;
            for (long long e = o1 + go1[gp]; e < o1 + go1[gp + 1]; e++) {
                const int2 en = TA.e1[e];
                const double *__restrict__ row = a.m2sq + (long long)((en.y >> 18) & 4095) * m1;
                AT cross = (AT)0;
                for (uint64_t z = mb; z;) {
                    double r8[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        r8[j] = 0.0;
                        if (z) {
                            r8[j] = row[tri32(jci_ctz(z), jci_ctz(z))];
                            z &= z - 1;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 8; j++) cross += (AT)r8[j];
                }
                AT v = (AT)TA.v1[e] + cross;
                if (en.y & (1 << 12)) v = -v;
                long long J = base + (long long)en.x * nq;
#pragma unroll
                for (int k = 0; k < K; k++) acc[k] += v * (AT)xT[J * K + k];
            }
        }
    }
    // add into the output: with a transposed descriptor the original (alpha-major)
    // layout's determinant (alpha ib_loc, beta ia_loc), else this determinant itself
    const long long O = a.tout ? L.boff[b] + ib_loc * L.gnA[p] + ia_loc
                               : L.boff[b] + ia_loc * nq + ib_loc;
    if (a.ldy) {
#pragma unroll
        for (int k = 0; k < K; k++) y[k * a.ldy + O] += (double)acc[k];
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) y[O * K + k] += (double)acc[k];
    }
}

// x[N][K] -> xT[N][K] with every block transposed (alpha-major -> beta-major), 16 x 16
// element tiles through LDS; tt_off[b] = first tile of block b
template <typename XT, int K>
__global__ __launch_bounds__(256) void k_block_transpose(const SpaceDesc *__restrict__ d,
                                                         const long long *__restrict__ tt_off,
                                                         const XT *__restrict__ x,
                                                         XT *__restrict__ xT)
{
    __shared__ XT t[16][17][K];
    const long long tile = blockIdx.x;
    const int b = tile_block(tt_off, d->nblk, tile);
    const long long na = d->s[0].gn[d->bp[b]], nb = d->s[1].gn[d->bq[b]];
    const long long ntc = (nb + 15) / 16;
    const long long rel = tile - tt_off[b];
    const long long r0 = rel / ntc * 16, c0 = rel % ntc * 16;
    const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
    const long long off = d->boff[b];
    if (r0 + ty < na && c0 + tx < nb) {
        const XT *src = x + (off + (r0 + ty) * nb + c0 + tx) * K;
#pragma unroll
        for (int k = 0; k < K; k++) t[ty][tx][k] = src[k];
    }
    __syncthreads();
    if (c0 + ty < nb && r0 + tx < na) {
        XT *dst = xT + (off + (c0 + ty) * na + r0 + tx) * K;
#pragma unroll
        for (int k = 0; k < K; k++) dst[k] = t[tx][ty][k];
    }
}

template <typename AT, typename XT, int W, int A, bool BSIDE>
static int launch_fast_a(const FastArgs &a, int nvec, const XT *x, double *y, hipStream_t st)
{
    size_t shm = sizeof(AT) * (size_t)A * (size_t)a.m1;
    dim3 g(xcd_grid(a.ntiles, a.xg)), b(W);
    switch (nvec) {
    case 1: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 1, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 2: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 2, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 3: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 3, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 4: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 4, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 5: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 5, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 6: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 6, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 7: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 7, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    case 8: hipLaunchKernelGGL((k_sigma_fast<AT, XT, 8, W, A, BSIDE>), g, b, shm, st, a, x, y); break;
    default: jci_set_error("sigma_fast: nvec must be 1..8"); return -1;
    }
    LAUNCH_CHECK2();
    return 0;
}

template <typename AT, typename XT, int W, bool BSIDE>
static int launch_fast(const FastArgs &a, int nvec, const XT *x, double *y, hipStream_t st)
{
    const int A = a.A > 0 ? a.A : fast_chunk(a.m1, sizeof(AT));
    if (A == 8) return launch_fast_a<AT, XT, W, 8, BSIDE>(a, nvec, x, y, st);
    if (A == 4) return launch_fast_a<AT, XT, W, 4, BSIDE>(a, nvec, x, y, st);
    if (A == 2) return launch_fast_a<AT, XT, W, 2, BSIDE>(a, nvec, x, y, st);
    if (A == 1) return launch_fast_a<AT, XT, W, 1, BSIDE>(a, nvec, x, y, st);
    jci_set_error("sigma_fast: alpha chunk %d not available (1, 2, 4 or 8)", A);
    return -1;
}

template <bool BSIDE>
static int fast_main(int prec, int nvec, const FastArgs &a, const void *x, double *y, hipStream_t st)
{
    if (a.ntiles <= 0) return 0;
    if (a.W == 64) {
        if (prec == 64) return launch_fast<double, double, 64, BSIDE>(a, nvec, (const double *)x, y, st);
        return launch_fast<float, float, 64, BSIDE>(a, nvec, (const float *)x, y, st);
    }
    if (a.W == 128) {
        if (prec == 64) return launch_fast<double, double, 128, BSIDE>(a, nvec, (const double *)x, y, st);
        return launch_fast<float, float, 128, BSIDE>(a, nvec, (const float *)x, y, st);
    }
    if (a.W == 256) {
        if (prec == 64) return launch_fast<double, double, 256, BSIDE>(a, nvec, (const double *)x, y, st);
        return launch_fast<float, float, 256, BSIDE>(a, nvec, (const float *)x, y, st);
    }
    if (prec == 64) return launch_fast<double, double, 512, BSIDE>(a, nvec, (const double *)x, y, st);
    return launch_fast<float, float, 512, BSIDE>(a, nvec, (const float *)x, y, st);
}

int jk_sigma_fast(int prec, int nvec, const FastArgs &a, const void *x, double *y, hipStream_t st)
{
    return fast_main<true>(prec, nvec, a, x, y, st);
}

template <typename XT, int K>
static int launch_transpose(const SpaceDesc *d, const long long *tt_off, long long ntt,
                            const XT *x, XT *xT, hipStream_t st)
{
    hipLaunchKernelGGL((k_block_transpose<XT, K>), dim3((unsigned)ntt), dim3(256), 0, st, d,
                       tt_off, x, xT);
    LAUNCH_CHECK2();
    return 0;
}

template <typename XT>
static int transpose_blocks(int nvec, const SpaceDesc *d, const long long *tt_off, long long ntt,
                            const XT *x, XT *xT, hipStream_t st)
{
    switch (nvec) {
    case 1: return launch_transpose<XT, 1>(d, tt_off, ntt, x, xT, st);
    case 2: return launch_transpose<XT, 2>(d, tt_off, ntt, x, xT, st);
    case 3: return launch_transpose<XT, 3>(d, tt_off, ntt, x, xT, st);
    case 4: return launch_transpose<XT, 4>(d, tt_off, ntt, x, xT, st);
    case 5: return launch_transpose<XT, 5>(d, tt_off, ntt, x, xT, st);
    case 6: return launch_transpose<XT, 6>(d, tt_off, ntt, x, xT, st);
    case 7: return launch_transpose<XT, 7>(d, tt_off, ntt, x, xT, st);
    case 8: return launch_transpose<XT, 8>(d, tt_off, ntt, x, xT, st);
    }
    jci_set_error("sigma_fast: nvec must be 1..8");
    return -1;
}

template <typename AT, typename XT, int W>
static int launch_beta(const FastArgs &t, int nvec, const XT *xT, double *y, hipStream_t st)
{
    dim3 g(xcd_grid(t.ntiles, t.xg)), b(W);
    switch (nvec) {
    case 1: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 1, W>), g, b, 0, st, t, xT, y); break;
    case 2: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 2, W>), g, b, 0, st, t, xT, y); break;
    case 3: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 3, W>), g, b, 0, st, t, xT, y); break;
    case 4: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 4, W>), g, b, 0, st, t, xT, y); break;
    case 5: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 5, W>), g, b, 0, st, t, xT, y); break;
    case 6: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 6, W>), g, b, 0, st, t, xT, y); break;
    case 7: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 7, W>), g, b, 0, st, t, xT, y); break;
    case 8: hipLaunchKernelGGL((k_sigma_fast_beta<AT, XT, 8, W>), g, b, 0, st, t, xT, y); break;
    default: jci_set_error("sigma_fast: nvec must be 1..8"); return -1;
    }
    LAUNCH_CHECK2();
    return 0;
}

#define RCHECK_SF(x)            \
    do {                        \
        int r_ = (x);           \
        if (r_) return r_;      \
    } while (0)

template <typename T>
static int beta_pass(const FastArgs &t, int nvec, const T *xT, double *y, hipStream_t st)
{
    if (t.ntiles <= 0) return 0;
    if (t.W == 64) return launch_beta<T, T, 64>(t, nvec, xT, y, st);
    if (t.W == 128) return launch_beta<T, T, 128>(t, nvec, xT, y, st);
    if (t.W == 256) return launch_beta<T, T, 256>(t, nvec, xT, y, st);
    return launch_beta<T, T, 512>(t, nvec, xT, y, st);
}

int jk_sigma_fast_multi(int prec, int nvec, const FastArgs *a, int na, const void *x, double *y,
                        hipStream_t st)
{
    for (int i = 0; i < na; i++) RCHECK_SF(fast_main<true>(prec, nvec, a[i], x, y, st));
    return 0;
}

int jk_sigma_fast_split(int prec, int nvec, const SpaceDesc *dD, const FastArgs *a, int na,
                        const FastArgs *t, int nt, const long long *tt_off, long long ntt,
                        const void *x, void *xT, double *y, hipStream_t st)
{
    // the block-transposed input first; then the main passes write y (each launch its own
    // blocks) and the complements add to it, each reading x or xT as its descriptor needs
    if (prec == 64) {
        RCHECK_SF(transpose_blocks<double>(nvec, dD, tt_off, ntt, (const double *)x, (double *)xT, st));
        for (int i = 0; i < na; i++)
            RCHECK_SF(fast_main<false>(prec, nvec, a[i], a[i].src_t ? xT : x, y, st));
        for (int i = 0; i < nt; i++)
            RCHECK_SF(beta_pass<double>(t[i], nvec, (const double *)(t[i].src_t ? xT : x), y, st));
        return 0;
    }
    RCHECK_SF(transpose_blocks<float>(nvec, dD, tt_off, ntt, (const float *)x, (float *)xT, st));
    for (int i = 0; i < na; i++)
        RCHECK_SF(fast_main<false>(prec, nvec, a[i], a[i].src_t ? xT : x, y, st));
    for (int i = 0; i < nt; i++)
        RCHECK_SF(beta_pass<float>(t[i], nvec, (const float *)(t[i].src_t ? xT : x), y, st));
    return 0;
}

// ---- single-block spaces (FCI) ------------------------------------------------------------
// One work-group of 256 threads owns (Ia, Ib_0 .. Ib_0+255).  The packed two-electron
// integrals live in LDS (n <= 16 orbitals), the work-group's alpha singles are decoded
// once into LDS, determinant addresses are Ja * Nb + Jb.  The alpha-beta doubles run with
// the lane's beta singles outside and the uniform alpha singles inside, so the hot loop
// has no barrier and reads no table from global memory.
struct FciAlpha {
    int ja;      // target alpha string
    int P;       // pair index tri(i, f)
    int Ptri;    // P(P+1)/2
    int neg;     // sign bit
};

template <typename AT, typename XT, int K>
__global__ __launch_bounds__(256) void k_sigma_fci(FciArgs a, const XT *__restrict__ x,
                                                   double *__restrict__ y)
{
    extern __shared__ unsigned char smem_f[];
    AT *m2s = (AT *)smem_f;
    FciAlpha *al = (FciAlpha *)(m2s + a.m2len_pad);
    AT *alv = (AT *)(al + FCI_MAXSINGLES);
    AT *Fa = alv + FCI_MAXSINGLES;
    const int tid = threadIdx.x;
    const long long Nb = a.nb;
    const long long nch = (Nb + 255) / 256;
    const long long Ia = blockIdx.x / nch;
    const long long ib = (blockIdx.x - Ia * nch) * 256 + tid;
    const bool active = ib < Nb;
    const uint64_t ma = a.ma[Ia];
    const uint64_t mb = a.mb[active ? ib : 0];
    for (int t = tid; t < a.m2len; t += 256) m2s[t] = (AT)a.m2[t];
    const long long s0 = a.ta.off1[Ia];
    const int nsa = (int)(a.ta.off1[Ia + 1] - s0);
    for (int t = tid; t < nsa; t += 256) {
        int2 en = a.ta.e1[s0 + t];
        int i = en.y & 63, f = (en.y >> 6) & 63;
        int P = (int)jci_tri(i, f);
        al[t] = FciAlpha{en.x, P, P * (P + 1) / 2, (en.y >> 12) & 1};
        alv[t] = (AT)a.ta.v1[s0 + t];
    }
    __syncthreads();
    for (int P = tid; P < a.m1; P += 256) {
        AT s = (AT)0;
        for (uint64_t z = ma; z; z &= z - 1) {
            int k = jci_ctz(z);
            int Q = k * (k + 1) / 2 + k;
            s += m2s[jci_tri(Q, P)];
        }
        Fa[P] = s;
    }
    __syncthreads();
    if (!active) return;
    const long long I = Ia * Nb + ib;
    AT acc[K];
    {
        AT h = (AT)a.hd[I];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = h * (AT)x[I * K + k];
    }
    const long long rowI = Ia * Nb;
    // beta singles and doubles of ib
    const long long b1s = a.tb.off1[ib], b1e = a.tb.off1[ib + 1];
    for (long long e = b1s; e < b1e; e++) {
        int2 en = a.tb.e1[e];
        int i = en.y & 63, f = (en.y >> 6) & 63;
        AT v = (AT)a.tb.v1[e] + Fa[jci_tri(i, f)];
        if (en.y & (1 << 12)) v = -v;
        long long J = rowI + en.x;
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
    }
    for (long long e = a.tb.off2[ib]; e < a.tb.off2[ib + 1]; e++) {
        AT v = (AT)a.tb.v2[e];
        long long J = rowI + a.tb.e2[e].x;
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
    }
    // alpha doubles (uniform)
    for (long long e = a.ta.off2[Ia]; e < a.ta.off2[Ia + 1]; e++) {
        AT v = (AT)a.ta.v2[e];
        long long J = (long long)a.ta.e2[e].x * Nb + ib;
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
    }
    // alpha singles: own element (cross-spin term over the beta electrons)
    int occb[16];
    int nob = 0;
    for (uint64_t z = mb; z; z &= z - 1) {
        int k = jci_ctz(z);
        occb[nob++] = k * (k + 1) / 2 + k;
    }
    for (int t = 0; t < nsa; t++) {
        FciAlpha A = al[t];
        AT cross = (AT)0;
        for (int u = 0; u < nob; u++) {
            int Q = occb[u];
            cross += m2s[Q >= A.P ? Q * (Q + 1) / 2 + A.P : A.Ptri + Q];
        }
        AT v = alv[t] + cross;
        if (A.neg) v = -v;
        long long J = (long long)A.ja * Nb + ib;
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
    }
    // alpha-beta doubles
    for (long long e = b1s; e < b1e; e++) {
        int2 en = a.tb.e1[e];
        int i = en.y & 63, f = (en.y >> 6) & 63;
        int Pb = (int)jci_tri(i, f);
        int Pbtri = Pb * (Pb + 1) / 2;
        int nb_ = (en.y >> 12) & 1;
        long long jb = en.x;
        for (int t = 0; t < nsa; t++) {
            FciAlpha A = al[t];
            AT v = m2s[A.P >= Pb ? A.Ptri + Pb : Pbtri + A.P];
            if (A.neg ^ nb_) v = -v;
            long long J = (long long)A.ja * Nb + jb;
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += v * (AT)x[J * K + k];
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) y[I * K + k] = (double)acc[k];
}

template <typename AT, typename XT>
static int launch_fci(const FciArgs &a, int nvec, const XT *x, double *y, hipStream_t st)
{
    size_t shm = sizeof(AT) * (size_t)a.m2len_pad + sizeof(FciAlpha) * FCI_MAXSINGLES +
                 sizeof(AT) * (FCI_MAXSINGLES + (size_t)a.m1);
    long long nch = (a.nb + 255) / 256;
    dim3 g((unsigned)(a.na * nch)), b(256);
    switch (nvec) {
    case 1: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 1>), g, b, shm, st, a, x, y); break;
    case 2: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 2>), g, b, shm, st, a, x, y); break;
    case 3: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 3>), g, b, shm, st, a, x, y); break;
    case 4: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 4>), g, b, shm, st, a, x, y); break;
    case 5: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 5>), g, b, shm, st, a, x, y); break;
    case 6: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 6>), g, b, shm, st, a, x, y); break;
    case 7: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 7>), g, b, shm, st, a, x, y); break;
    case 8: hipLaunchKernelGGL((k_sigma_fci<AT, XT, 8>), g, b, shm, st, a, x, y); break;
    default: jci_set_error("sigma_fci: nvec must be 1..8"); return -1;
    }
    LAUNCH_CHECK2();
    return 0;
}

int jk_sigma_fci(int prec, int nvec, const FciArgs &a, const void *x, double *y, hipStream_t st)
{
    if (prec == 64) return launch_fci<double, double>(a, nvec, (const double *)x, y, st);
    return launch_fci<float, float>(a, nvec, (const float *)x, y, st);
}
