// Batched Davidson kernels: one launch serves every active CI instance.
//
// Each launch reads a header (work counts) and an array of per-instance descriptors
// (BDesc / BSigma) from device memory.  Work is flattened over instances: instance i
// owns flattened blocks [blk0, blk0 + nblk) and a work-group finds its instance by binary
// search over blk0.  The kernels are persistent: a fixed grid strides over the
// flattened blocks, so launch parameters never change between iterations (only the
// device-side header and descriptors do), which is what lets a step be recorded once as a
// HIP graph and replayed.
#include "jci_internal.h"
#include "slater.h"
#include "sigma_fast.h"
#include "batch.h"

#define BLAUNCH_CHECK()                                                  \
    do {                                                                 \
        hipError_t e_ = hipGetLastError();                               \
        if (e_ != hipSuccess) {                                          \
            jci_set_error("%s:%d launch: %s", __FILE__, __LINE__,        \
                          hipGetErrorString(e_));                        \
            return -2;                                                   \
        }                                                                \
    } while (0)

__device__ __forceinline__ int find_inst(const BDesc *__restrict__ d, int n, long long blk)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d[mid].blk0 <= blk) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---- Gram: out = A^T B (m x k), A and B vector-major with leading dimension ld ----------
// FP64 MFMA 16x16x4 on 32-row LDS tiles (operand layout as k_gram); partial per block.
// MAXM bounds m and k over the launch (checked by the host): the LDS tiles, the staging
// registers and the (MAXM/16)^2 output tiles spread over the four waves are sized to it.
#define BG_ROWS 32
template <int MAXM>
__global__ __launch_bounds__(256) void k_bgram(const BHeader *__restrict__ H,
                                               const BDesc *__restrict__ D,
                                               double *__restrict__ partial)
{
    constexpr int NT = ((MAXM / 16) * (MAXM / 16) + 3) / 4;   // output tiles per wave
    __shared__ double sA[BG_ROWS][MAXM + 1];
    __shared__ double sB[BG_ROWS][MAXM + 1];
    const int n = H->n;
    const long long total = H->total;
    for (long long blk = blockIdx.x; blk < total; blk += gridDim.x) {
    __syncthreads();
    const int inst = find_inst(D, n, blk);
    const BDesc &d = D[inst];
    const int b = (int)(blk - d.blk0);
    const int m = d.m, w = d.k;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mt = (m + 15) / 16, wt = (w + 15) / 16, ntile = mt * wt;
    const long long chunk = (d.N + d.nblk - 1) / d.nblk;
    const long long r0 = (long long)b * chunk;
    const long long r1 = r0 + chunk < d.N ? r0 + chunk : d.N;
    typedef double v4d __attribute__((ext_vector_type(4)));
    v4d acc[NT];
    for (int t = 0; t < NT; t++) acc[t] = (v4d){0.0, 0.0, 0.0, 0.0};
    // the chunk's values this thread stages (element idx = tid + 256 q: column idx / BG_ROWS,
    // row idx % BG_ROWS), loaded into registers one chunk ahead so that the loads of the
    // next chunk are in flight during the products of this one
    constexpr int QA = BG_ROWS * MAXM / 256, QB = BG_ROWS * MAXM / 256;
    double ra[QA], rb[QB];
    auto gload = [&](long long base) {
        const int rows = (int)((r1 - base) < BG_ROWS ? (r1 - base) : BG_ROWS);
#pragma unroll
        for (int q = 0; q < QA; q++) {
            const int idx = tid + 256 * q, c = idx / BG_ROWS, r = idx % BG_ROWS;
            ra[q] = (idx < BG_ROWS * mt * 16 && c < m && r < rows) ? d.A[(long long)c * d.ld + base + r] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int idx = tid + 256 * q, c = idx / BG_ROWS, r = idx % BG_ROWS;
            rb[q] = (idx < BG_ROWS * wt * 16 && c < w && r < rows) ? d.B[(long long)c * d.ld + base + r] : 0.0;
        }
    };
    if (r0 < r1) gload(r0);
    for (long long base = r0; base < r1; base += BG_ROWS) {
#pragma unroll
        for (int q = 0; q < QA; q++) {
            const int idx = tid + 256 * q;
            if (idx < BG_ROWS * mt * 16) sA[idx % BG_ROWS][idx / BG_ROWS] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int idx = tid + 256 * q;
            if (idx < BG_ROWS * wt * 16) sB[idx % BG_ROWS][idx / BG_ROWS] = rb[q];
        }
        __syncthreads();
        if (base + BG_ROWS < r1) gload(base + BG_ROWS);
        for (int t = 0; t < NT; t++) {
            int tile = wave + 4 * t;
            if (tile >= ntile) break;
            int ti = tile / wt, tj = tile % wt;
            for (int k0 = 0; k0 < BG_ROWS; k0 += 4) {
                double a = sA[k0 + lane / 16][ti * 16 + lane % 16];
                double bb = sB[k0 + lane / 16][tj * 16 + lane % 16];
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    double *out = partial + d.pofs + (long long)b * m * w;
    for (int t = 0; t < NT; t++) {
        int tile = wave + 4 * t;
        if (tile >= ntile) break;
        int ti = tile / wt, tj = tile % wt;
        for (int v = 0; v < 4; v++) {
            int i = ti * 16 + lane / 16 + 4 * v, j = tj * 16 + lane % 16;
            if (i < m && j < w) out[i * w + j] = acc[t][v];
        }
    }
    }
}

// fixed-order sum of the per-block partials of each instance: out = sum_b partial_b.
// blockIdx.x: instance; blockIdx.y: one of BPS_SPLIT blocks sharing its elements, one
// element per wave at a time (the order of each element's sum does not depend on the split)
#define BPS_SPLIT 16
__global__ void k_bpartial_sum(const BHeader *__restrict__ H, const BDesc *__restrict__ D,
                               const double *__restrict__ partial)
{
    if ((int)blockIdx.x >= H->n) return;
    const BDesc &d = D[blockIdx.x];
    int cnt = d.m * d.k, lane = threadIdx.x & 63;
    for (int e = (int)blockIdx.y * 4 + (threadIdx.x >> 6); e < cnt; e += 4 * (int)gridDim.y) {
        double s = 0.0;
        for (int b = lane; b < d.nblk; b += 64) s += partial[d.pofs + (long long)b * cnt + e];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) d.out[e] = s;
    }
}

int jb_gram(const BHeader *h, const BDesc *dD, int n_cap, int mmax, double *partial,
            hipStream_t st)
{
    if (mmax > JCI_MAXBASIS) return -1;
    if (mmax <= 16)
        hipLaunchKernelGGL(k_bgram<16>, dim3(JB_GRID), dim3(256), 0, st, h, dD, partial);
    else if (mmax <= 32)
        hipLaunchKernelGGL(k_bgram<32>, dim3(JB_GRID), dim3(256), 0, st, h, dD, partial);
    else if (mmax <= 48)
        hipLaunchKernelGGL(k_bgram<48>, dim3(JB_GRID), dim3(256), 0, st, h, dD, partial);
    else
        hipLaunchKernelGGL(k_bgram<JCI_MAXBASIS>, dim3(JB_GRID), dim3(256), 0, st, h, dD, partial);
    hipLaunchKernelGGL(k_bpartial_sum, dim3(n_cap, BPS_SPLIT), dim3(256), 0, st, h, dD,
                       (const double *)partial);
    BLAUNCH_CHECK();
    return 0;
}

// ---- combine: O_j = beta * O_j + sum_i A_i C[i*k + j] (in place when O == A) -----------
// MAXM bounds m over the launch (the host picks the smallest instance that holds it): a
// thread keeps the m inputs of its row in registers, so every input of the row is read before
// any output is written (in place is safe), and the register array is only as large as the
// launch needs.  The coefficients are staged in LDS when the instance changes; every
// descriptor has k <= m.
template <int MAXM>
__global__ __launch_bounds__(256) void k_bcombine(const BHeader *__restrict__ H,
                                                  const BDesc *__restrict__ D)
{
    __shared__ double sC[MAXM * MAXM];   // k <= m <= MAXM (checked by the host)
    const int n = H->n;
    const long long total = H->total;
    int cur = -1;
    for (long long blk = blockIdx.x; blk < total; blk += gridDim.x) {
        const int inst = find_inst(D, n, blk);
        const BDesc &d = D[inst];
        if (inst != cur) {
            __syncthreads();
            for (int t = threadIdx.x; t < d.m * d.k; t += blockDim.x) sC[t] = d.C[t];
            __syncthreads();
            cur = inst;
        }
        const long long r = (long long)(blk - d.blk0) * 256 + threadIdx.x;
        if (r >= d.N) continue;
        double a[MAXM];
#pragma unroll
        for (int i = 0; i < MAXM; i++)
            if (i < d.m) a[i] = d.A[(long long)i * d.ld + r];
        for (int j = 0; j < d.k; j++) {
            double s = d.beta == 0.0 ? 0.0 : d.beta * d.O[(long long)j * d.ld + r];
#pragma unroll
            for (int i = 0; i < MAXM; i++)
                if (i < d.m) s += a[i] * sC[i * d.k + j];
            d.O[(long long)j * d.ld + r] = s;
        }
    }
}

int jb_combine(const BHeader *h, const BDesc *dD, int mmax, hipStream_t st)
{
    if (mmax > JCI_MAXBASIS) return -1;
    if (mmax <= 16)
        hipLaunchKernelGGL(k_bcombine<16>, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    else if (mmax <= 32)
        hipLaunchKernelGGL(k_bcombine<32>, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    else if (mmax <= 48)
        hipLaunchKernelGGL(k_bcombine<48>, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    else
        hipLaunchKernelGGL(k_bcombine<JCI_MAXBASIS>, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    BLAUNCH_CHECK();
    return 0;
}

// ---- residuals + preconditioned corrections -------------------------------------------------
// R_r = sum_i C[i*k + r] (B_i - th_r A_i), O_r = -R_r / (hd - th_r); partial norms^2
__global__ __launch_bounds__(256) void k_bresidual(const BHeader *__restrict__ H,
                                                   const BDesc *__restrict__ D,
                                                   double *__restrict__ partial)
{
    __shared__ double sC[JCI_MAXBASIS * JCI_MAXK];
    __shared__ double sT[JCI_MAXK];
    __shared__ double red[JCI_MAXK][256];
    const int n = H->n;
    const long long total = H->total;
    for (long long blk = blockIdx.x; blk < total; blk += gridDim.x) {
    const int inst = find_inst(D, n, blk);
    const BDesc &d = D[inst];
    const int m = d.m, k = d.k;
    __syncthreads();
    for (int t = threadIdx.x; t < m * k; t += blockDim.x) sC[t] = d.C[t];
    if (threadIdx.x < k) sT[threadIdx.x] = d.th[threadIdx.x];
    __syncthreads();
    const int b = (int)(blk - d.blk0);
    const long long chunk = (d.N + d.nblk - 1) / d.nblk;
    const long long r0 = (long long)b * chunk, r1 = r0 + chunk < d.N ? r0 + chunk : d.N;
    double nrm[JCI_MAXK];
    for (int r = 0; r < k; r++) nrm[r] = 0.0;
    for (long long row = r0 + threadIdx.x; row < r1; row += 256) {
        double res[JCI_MAXK];
        for (int r = 0; r < k; r++) res[r] = 0.0;
        // four basis vectors per step: their eight loads are issued together
        int i = 0;
        for (; i + 4 <= m; i += 4) {
            double a[4], sv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                a[q] = d.A[(long long)(i + q) * d.ld + row];
                sv[q] = d.B[(long long)(i + q) * d.ld + row];
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                for (int r = 0; r < k; r++) res[r] += sC[(i + q) * k + r] * (sv[q] - sT[r] * a[q]);
        }
        for (; i < m; i++) {
            double a = d.A[(long long)i * d.ld + row], sv = d.B[(long long)i * d.ld + row];
            for (int r = 0; r < k; r++) res[r] += sC[i * k + r] * (sv - sT[r] * a);
        }
        double h = d.hd[row];
        for (int r = 0; r < k; r++) {
            nrm[r] += res[r] * res[r];
            double den = h - sT[r];
            if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
            d.O[(long long)r * d.ld + row] = -res[r] / den;
        }
    }
    for (int r = 0; r < k; r++) red[r][threadIdx.x] = nrm[r];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int r = 0; r < k; r++) red[r][threadIdx.x] += red[r][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < k) partial[d.pofs + (long long)b * k + threadIdx.x] = red[threadIdx.x][0];
    }
}

// sums of k entries per block (norms): out[r] = sum_b partial[b*k + r]
__global__ void k_bpartial_vec(const BHeader *__restrict__ H, const BDesc *__restrict__ D,
                               const double *__restrict__ partial)
{
    if ((int)blockIdx.x >= H->n) return;
    const BDesc &d = D[blockIdx.x];
    const int lane = threadIdx.x & 63;
    // one wave per entry: lanes stride over the blocks, then a fixed shuffle tree
    for (int r = threadIdx.x >> 6; r < d.k; r += (int)blockDim.x >> 6) {
        double s = 0.0;
        for (int b = lane; b < d.nblk; b += 64) s += partial[d.pofs + (long long)b * d.k + r];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) d.out[r] = s;
    }
}

int jb_residual(const BHeader *h, const BDesc *dD, int n_cap, double *partial, hipStream_t st)
{
    hipLaunchKernelGGL(k_bresidual, dim3(JB_GRID), dim3(256), 0, st, h, dD, partial);
    hipLaunchKernelGGL(k_bpartial_vec, dim3(n_cap), dim3(256), 0, st, h, dD,
                       (const double *)partial);
    BLAUNCH_CHECK();
    return 0;
}

// ---- pack / unpack between vector-major columns and determinant-major blocks ----------------
// pack:   O[r*k + j] = A[(m + j)*ld + r]     (k columns starting at column m)
// unpack: O[(m + j)*ld + r] = A[r*k + j]
template <bool PACK>
__global__ __launch_bounds__(256) void k_bpack(const BHeader *__restrict__ H,
                                               const BDesc *__restrict__ D)
{
    const int n = H->n;
    const long long total = H->total;
    for (long long blk = blockIdx.x; blk < total; blk += gridDim.x) {
        const int inst = find_inst(D, n, blk);
        const BDesc &d = D[inst];
        long long r = (long long)(blk - d.blk0) * 256 + threadIdx.x;
        if (r >= d.N) continue;
        for (int j = 0; j < d.k; j++) {
            if (PACK) d.O[r * d.k + j] = d.A[(long long)(d.m + j) * d.ld + r];
            else d.O[(long long)(d.m + j) * d.ld + r] = d.A[r * d.k + j];
        }
    }
}

int jb_pack(const BHeader *h, const BDesc *dD, int pack, hipStream_t st)
{
    if (pack) hipLaunchKernelGGL(k_bpack<true>, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    else hipLaunchKernelGGL(k_bpack<false>, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    BLAUNCH_CHECK();
    return 0;
}

// ---- batched matrix-free sigma -----------------------------------------------------------
// y = H x for every instance; x, y determinant-major [N][k] (k <= K); one work-group per
// (instance, alpha string, beta tile) as k_sigma_fast, with the instance found from the
// tile index.
__device__ __forceinline__ int find_sinst(const BSigma *__restrict__ S, int n, long long t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (S[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int tile_block2(const long long *__restrict__ toff, int nblk, long long t)
{
    int lo = 0, hi = nblk - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (toff[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename AT, typename XT, int K, int W>
__global__ __launch_bounds__(W) void k_bsigma(const BHeader *__restrict__ H,
                                              const BSigma *__restrict__ S)
{
    extern __shared__ unsigned char smem_b[];
    const int n = H->n;
    const long long total = H->total;
    for (long long tile_g = blockIdx.x; tile_g < total; tile_g += gridDim.x) {
    __syncthreads();
    const int inst = find_sinst(S, n, tile_g);
    const BSigma &bs = S[inst];
    const FastArgs &a = bs.a;
    const XT *__restrict__ x = (const XT *)bs.x;
    double *__restrict__ y = bs.y;
    const int kk = bs.k;
    AT *g = (AT *)smem_b;
    AT *Fa = g + a.m1;
    const SpaceDesc *__restrict__ d = a.d;
    const int tid = threadIdx.x;
    const long long tile = tile_g - bs.tile0;
    const int b = tile_block2(a.tile_off, d->nblk, tile);
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
#pragma unroll
    for (int v = 0; v < K; v++) acc[v] = (AT)0;
    if (active) {
        AT h = (AT)a.hd[I];
#pragma unroll
        for (int v = 0; v < K; v++)
            if (v < kk) acc[v] = h * (AT)x[I * kk + v];
    }
    const SpinDesc &SA = d->s[0], &SB = d->s[1];
    const ExcView &TA = a.ta, &TB = a.tb;
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
            for (int u = 0; u < K; u++)
                if (u < kk) acc[u] += v * (AT)x[J * kk + u];
        }
        for (long long e = TB.off2[Ib]; e < TB.off2[Ib + 1]; e++) {
            int2 en = TB.e2[e];
            int gq = (en.y >> 25) & 31;
            int bk = d->blkid[p][gq];
            if (bk < 0) continue;
            long long J = d->boff[bk] + ia_loc * SB.gn[gq] + (en.x - SB.gstart[gq]);
            AT v = (AT)TB.v2[e];
#pragma unroll
            for (int u = 0; u < K; u++)
                if (u < kk) acc[u] += v * (AT)x[J * kk + u];
        }
    }
    for (long long e = TA.off2[Ia]; e < TA.off2[Ia + 1]; e++) {
        int2 en = TA.e2[e];
        int gp = (en.y >> 25) & 31;
        int bk = d->blkid[gp][q];
        if (bk < 0 || !active) continue;
        long long J = d->boff[bk] + (en.x - SA.gstart[gp]) * nq + ib_loc;
        AT v = (AT)TA.v2[e];
#pragma unroll
        for (int u = 0; u < K; u++)
            if (u < kk) acc[u] += v * (AT)x[J * kk + u];
    }
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
            for (int u = 0; u < K; u++)
                if (u < kk) acc[u] += v * (AT)x[J * kk + u];
        }
        for (long long e2 = b1; e2 < e1end; e2++) {
            int2 eb = TB.e1[e2];
            int gq = (eb.y >> 13) & 31;
            int bk2 = d->blkid[gp][gq];
            if (bk2 < 0) continue;
            int k2 = eb.y & 63, l2 = (eb.y >> 6) & 63;
            AT v = g[jci_tri(k2, l2)];
            if (neg != (bool)(eb.y & (1 << 12))) v = -v;
            long long J = d->boff[bk2] + ja_loc * SB.gn[gq] + (eb.x - SB.gstart[gq]);
#pragma unroll
            for (int u = 0; u < K; u++)
                if (u < kk) acc[u] += v * (AT)x[J * kk + u];
        }
    }
    if (active) {
#pragma unroll
        for (int v = 0; v < K; v++)
            if (v < kk) y[I * kk + v] = (double)acc[v];
    }
    }
}

template <typename AT, typename XT, int W>
static int bsigma_launch(int K, const BHeader *h, const BSigma *dS, int m1max, hipStream_t st)
{
    size_t shm = sizeof(AT) * 2 * (size_t)m1max;
    dim3 g(JB_SGRID), b(W);
    if (K <= 1) hipLaunchKernelGGL((k_bsigma<AT, XT, 1, W>), g, b, shm, st, h, dS);
    else if (K <= 2) hipLaunchKernelGGL((k_bsigma<AT, XT, 2, W>), g, b, shm, st, h, dS);
    else if (K <= 4) hipLaunchKernelGGL((k_bsigma<AT, XT, 4, W>), g, b, shm, st, h, dS);
    else hipLaunchKernelGGL((k_bsigma<AT, XT, 8, W>), g, b, shm, st, h, dS);
    BLAUNCH_CHECK();
    return 0;
}

int jb_sigma(int prec, int W, int K, const BHeader *h, const BSigma *dS, int m1max,
             hipStream_t st)
{
    if (W == 64)
        return prec == 64 ? bsigma_launch<double, double, 64>(K, h, dS, m1max, st)
                          : bsigma_launch<float, float, 64>(K, h, dS, m1max, st);
    if (W == 128)
        return prec == 64 ? bsigma_launch<double, double, 128>(K, h, dS, m1max, st)
                          : bsigma_launch<float, float, 128>(K, h, dS, m1max, st);
    return prec == 64 ? bsigma_launch<double, double, 256>(K, h, dS, m1max, st)
                      : bsigma_launch<float, float, 256>(K, h, dS, m1max, st);
}

// FP64 -> FP32 copy of every instance's packed block (N*k entries)
// (flattened blocks cover N*k entries per instance)
__global__ __launch_bounds__(256) void k_bto32(const BHeader *__restrict__ H,
                                               const BDesc *__restrict__ D)
{
    const int n = H->n;
    const long long total = H->total;
    for (long long blk = blockIdx.x; blk < total; blk += gridDim.x) {
        const int inst = find_inst(D, n, blk);
        const BDesc &d = D[inst];
        long long r = (long long)(blk - d.blk0) * 256 + threadIdx.x;
        if (r < d.N * d.k) ((float *)d.O)[r] = (float)d.A[r];
    }
}

int jb_to32(const BHeader *h, const BDesc *dD, hipStream_t st)
{
    hipLaunchKernelGGL(k_bto32, dim3(JB_GRID), dim3(256), 0, st, h, dD);
    BLAUNCH_CHECK();
    return 0;
}
