// Dense-string full-CI sigma on MFMA (see sigma_gemm.h for the decomposition).
#include <hip/hip_runtime.h>
#include <algorithm>
#include "sigma_gemm.h"
#include "slater.h"

#define RCHECK(x)                      \
    do {                               \
        int rc_ = (x);                 \
        if (rc_) return rc_;           \
    } while (0)

#define GF_CHECK()                                                       \
    do {                                                                 \
        hipError_t e_ = hipGetLastError();                               \
        if (e_ != hipSuccess) {                                          \
            jci_set_error("%s:%d launch: %s", __FILE__, __LINE__,        \
                          hipGetErrorString(e_));                        \
            return -2;                                                   \
        }                                                                \
    } while (0)

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// device-memory (global address space) views of pointers read from descriptors, so their
// loads are global loads (vector-memory counter only) rather than flat ones, which a later
// LDS wait would also have to drain
template <typename T> __device__ __forceinline__ const __attribute__((address_space(1))) T *
gptr(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T> __device__ __forceinline__ __attribute__((address_space(1))) T *gptr_w(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// 16x16 output tile, 4-deep reduction per instruction.  Operand lanes: A[l % 16][l / 16],
// B[l / 16][l % 16]; accumulator element v of lane l sits in column l % 16 and row row(l, v).
template <typename T> struct Mf;
template <> struct Mf<double> {
    typedef v4d acc;
    static __device__ __forceinline__ acc zero() { return (v4d){0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc mma(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int v) { return l / 16 + 4 * v; }
};
template <> struct Mf<float> {
    typedef v4f acc;
    static __device__ __forceinline__ acc zero() { return (v4f){0.0f, 0.0f, 0.0f, 0.0f}; }
    static __device__ __forceinline__ acc mma(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int v) { return 4 * (l / 16) + v; }
};

// ---- set-up kernels ---------------------------------------------------------------------------
__global__ void k_gf_dense(long long nstr, const uint64_t *__restrict__ mask,
                           const long long *__restrict__ off1, const int2 *__restrict__ e1,
                           const double *__restrict__ v1, const long long *__restrict__ off2,
                           const int2 *__restrict__ e2, const double *__restrict__ v2,
                           const double *__restrict__ m1, const double *__restrict__ m2,
                           double *__restrict__ H)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nstr) return;
    const uint64_t s = mask[I];
    double d = 0.0;
    for (uint64_t x = s; x; x &= x - 1) {
        int i = jci_ctz(x);
        d += jci_h1(m1, i, i);
        for (uint64_t y = s & ((1ULL << i) - 1ULL); y; y &= y - 1) {
            int j = jci_ctz(y);
            d += jci_eri(m2, i, i, j, j) - jci_eri(m2, i, j, i, j);
        }
    }
    double *row = H + I * nstr;
    row[I] = d;
    for (long long e = off1[I]; e < off1[I + 1]; e++) {
        int2 en = e1[e];
        double v = v1[e];
        row[en.x] = (en.y & (1 << 12)) ? -v : v;
    }
    for (long long e = off2[I]; e < off2[I + 1]; e++) row[e2[e].x] = v2[e];
}

__global__ void k_gf_lists(long long nstr, const uint64_t *__restrict__ mask,
                           const long long *__restrict__ off1, const int2 *__restrict__ e1, int L,
                           int diag_one, uint32_t *__restrict__ out)
{
    long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nstr) return;
    uint32_t *o = out + I * L;
    int n = 0;
    for (long long e = off1[I]; e < off1[I + 1] && n < L; e++) {
        int2 en = e1[e];
        int i = en.y & 63, f = (en.y >> 6) & 63;
        uint32_t P = (uint32_t)jci_tri(i, f);
        o[n++] = (uint32_t)en.x | (P << 16) | ((uint32_t)((en.y >> 12) & 1) << 31);
    }
    if (diag_one) {
        if (mask[I] && n < L) o[n++] = (uint32_t)I | ((uint32_t)GF_PDIAG << 16);
        return;
    }
    for (uint64_t x = mask[I]; x && n < L; x &= x - 1) {
        int k = jci_ctz(x);
        o[n++] = (uint32_t)I | ((uint32_t)jci_tri(k, k) << 16);
    }
}

// E row of each alpha-list entry: the position of the entry's pair in its target's list
__global__ void k_gf_rank(long long nstr, int L, const uint32_t *__restrict__ in,
                          uint32_t *__restrict__ out, int *__restrict__ missing)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nstr * L) return;
    const uint32_t w = in[idx];
    const long long J = w & 0xFFFFu;
    const uint32_t P = (w >> 16) & 0xFFu;
    uint32_t pos = 0;
    bool found = false;
    if (J < nstr)
        for (int e = 0; e < L; e++)
            if (((in[J * L + e] >> 16) & 0xFFu) == P) { pos = (uint32_t)e; found = true; break; }
    if (!found) atomicAdd(missing, 1);
    out[idx] = (w & ~(0x7Fu << 24)) | (pos << 24);
}

__global__ void k_gf_diagrows(long long nstr, int MP, const uint64_t *__restrict__ mask,
                              const double *__restrict__ M64, double *__restrict__ o64,
                              float *__restrict__ o32)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nstr * MP) return;
    const long long J = idx / MP;
    const int Q = (int)(idx - J * MP);
    double s = 0.0;
    for (uint64_t x = mask[J]; x; x &= x - 1) {
        const int k = jci_ctz(x);
        s += M64[jci_tri(k, k) * MP + Q];
    }
    o64[idx] = s;
    o32[idx] = (float)s;
}

__global__ void k_gf_m2pad(int m1, int MP, const double *__restrict__ sq, double *__restrict__ o64,
                           float *__restrict__ o32)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= MP * MP) return;
    int P = idx / MP, Q = idx - P * MP;
    double v = (P < m1 && Q < m1) ? sq[(long long)P * m1 + Q] : 0.0;
    o64[idx] = v;
    o32[idx] = (float)v;
}


// ---- grouped kernels ------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ const T *hmat(const GemmFci &g, int spin);
template <> __device__ __forceinline__ const double *hmat<double>(const GemmFci &g, int spin)
{
    return g.H64[spin];
}
template <> __device__ __forceinline__ const float *hmat<float>(const GemmFci &g, int spin)
{
    return g.H32[spin];
}
template <typename T> __device__ __forceinline__ const T *mpad(const GemmFci &g);
template <> __device__ __forceinline__ const double *mpad<double>(const GemmFci &g) { return g.M64; }
template <> __device__ __forceinline__ const float *mpad<float>(const GemmFci &g) { return g.M32; }
template <typename T> __device__ __forceinline__ const T *drows(const GemmFci &g);
template <> __device__ __forceinline__ const double *drows<double>(const GemmFci &g) { return g.D64; }
template <> __device__ __forceinline__ const float *drows<float>(const GemmFci &g) { return g.D32; }

// input vector k of a task in the arithmetic type
template <typename T> __device__ __forceinline__ const T *xin(const GfTask &t, int k);
template <> __device__ __forceinline__ const double *xin<double>(const GfTask &t, int k)
{
    return t.x64 + (long long)k * t.ldx64;
}
template <> __device__ __forceinline__ const float *xin<float>(const GfTask &t, int k)
{
    return t.x32 + (long long)k * t.g.na * t.g.nb;
}

// task owning flattened unit u of kernel `which` (tasks ordered by base)
__device__ __forceinline__ int find_task(const GfTask *__restrict__ T, int n, long long u, int which)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (T[mid].base[which] <= u) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// FP64 -> FP32 copies of the input vectors, 256 elements per unit
__global__ __launch_bounds__(256) void k_gcv(const GfHeader *__restrict__ H,
                                             const GfTask *__restrict__ tasks)
{
    const int n = H->n;
    const long long total = H->total[GF_CV];
    for (long long u = blockIdx.x; u < total; u += gridDim.x) {
        const GfTask &t = tasks[find_task(tasks, n, u, GF_CV)];
        const long long N = t.g.na * t.g.nb;
        const long long idx = (u - t.base[GF_CV]) * 256 + threadIdx.x;
        if (idx < N * t.K) {
            const long long k = idx / N, i = idx - k * N;
            t.x32[idx] = (float)t.x64[k * t.ldx64 + i];
        }
    }
}

// C(64x64 tile at m0, n0) += A B over Kr, A row-major (lda), B row-major (ldb), M x N bounds.
// Two LDS buffers: the global loads of step k0 + 16 are issued before the products of step
// k0 and stored into the other buffer after them, one barrier per step.
template <typename T>
__device__ __forceinline__ void mma_tile(typename Mf<T>::acc (&acc)[2][2], const T *__restrict__ A,
                                         long long lda, const T *__restrict__ B, long long ldb,
                                         long long M, long long N, long long Kr, long long m0,
                                         long long n0, T (*As)[16][65], T (*Bs)[16][65])
{
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int ar = tid >> 2, ak = (tid & 3) * 4;
    const int bk = tid >> 4, bc = (tid & 15) * 4;
    const long long gm = m0 + ar;
    const auto Ag = gptr(A), Bg = gptr(B);
    T ra[4], rb[4];
    auto gload = [&](long long k0) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long long gk = k0 + ak + j;
            ra[j] = (gm < M && gk < Kr) ? Ag[gm * lda + gk] : (T)0;
        }
        const long long gk = k0 + bk;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long long gn = n0 + bc + j;
            rb[j] = (gk < Kr && gn < N) ? Bg[gk * ldb + gn] : (T)0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            As[buf][ak + j][ar] = ra[j];
            Bs[buf][bk][bc + j] = rb[j];
        }
    };
    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (long long k0 = 0; k0 < Kr; k0 += 16) {
        const bool more = k0 + 16 < Kr;
        if (more) gload(k0 + 16);
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kr = ks * 4 + l / 16;
            const T a0 = As[buf][kr][wm * 32 + l % 16], a1 = As[buf][kr][wm * 32 + 16 + l % 16];
            const T b0 = Bs[buf][kr][wn * 32 + l % 16], b1 = Bs[buf][kr][wn * 32 + 16 + l % 16];
            acc[0][0] = Mf<T>::mma(a0, b0, acc[0][0]);
            acc[0][1] = Mf<T>::mma(a0, b1, acc[0][1]);
            acc[1][0] = Mf<T>::mma(a1, b0, acc[1][0]);
            acc[1][1] = Mf<T>::mma(a1, b1, acc[1][1]);
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
}

// same-spin parts: y_k = Haa X_k + X_k Hbb on 64 x 64 tiles (k, alpha rows, beta columns).
// Four blocks per CU (at most 128 registers, 33 KB of LDS in FP64): the 1024 tiles of a
// 2002 x 2002 string space are then resident at once instead of running as 768 + 256.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_gss(const GfHeader *__restrict__ H,
                                             const GfTask *__restrict__ tasks)
{
    __shared__ T As[2][16][65];
    __shared__ T Bs[2][16][65];
    const int n = H->n;
    const long long total = H->total[GF_SS];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w >> 1, wn = w & 1;
    for (long long u = blockIdx.x; u < total; u += gridDim.x) {
        const GfTask &t = tasks[find_task(tasks, n, u, GF_SS)];
        const GemmFci &g = t.g;
        const long long na = g.na, nb = g.nb;
        const long long tn = (nb + 63) / 64, tm = (na + 63) / 64;
        const long long loc = u - t.base[GF_SS];
        const int k = (int)(loc / (tm * tn));
        const long long rem = loc - (long long)k * tm * tn;
        const long long m0 = (rem / tn) * 64, n0 = (rem % tn) * 64;
        const T *X = xin<T>(t, k);
        typename Mf<T>::acc acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = Mf<T>::zero();
        mma_tile<T>(acc, hmat<T>(g, 0), na, X, nb, na, nb, na, m0, n0, As, Bs);
        mma_tile<T>(acc, X, nb, hmat<T>(g, 1), nb, na, nb, nb, m0, n0, As, Bs);
        double *y = t.y + (long long)k * t.ldy;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const long long r = m0 + wm * 32 + i * 16 + Mf<T>::row(l, v);
                    const long long c = n0 + wn * 32 + j * 16 + l % 16;
                    if (r < na && c < nb) y[r * nb + c] = (double)acc[i][j][v];
                }
    }
}

// The D tile sits in LDS column-major, Dt[column][Q], and the reduction index of the
// products is permuted: step ks, lane group kk = l / 16 takes Q = kk * KS + ks (the (P|Q)
// fragments use the same order).  A lane's B values for VK consecutive steps are then
// adjacent and come in one 16-byte LDS read.  Row stride GF_LDT elements (a multiple of
// 16 bytes, 4 banks past a multiple of 64 banks).
template <typename T> struct GeL;
template <> struct GeL<double> {
    static constexpr int LDT = GF_MAXMP + 2;
    static constexpr int VK = 2;
    typedef double2 vec;
    static __device__ __forceinline__ double at(const vec &v, int q) { return q ? v.y : v.x; }
};
template <> struct GeL<float> {
    static constexpr int LDT = GF_MAXMP + 4;
    static constexpr int VK = 4;
    typedef float4 vec;
    static __device__ __forceinline__ float at(const vec &v, int q)
    {
        return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
    }
};

// Products of one D tile for the wave's row tile: acc[j] += A . D(column tile j) over the
// KS reduction steps, straight-line code, the B vectors of the next group of steps read
// while the products of the current group run (group gi = steps gi*VK .. gi*VK + VK - 1).
template <typename T, int BN, int KS>
__device__ __forceinline__ void gE_mma(typename Mf<T>::acc (&acc)[BN / 16],
                                       const T (&a)[GF_MAXMP / 4], const T (*Dt)[GeL<T>::LDT],
                                       int l)
{
    constexpr int NCT = BN / 16, VK = GeL<T>::VK, NG = KS / VK;
    typedef typename GeL<T>::vec V;
    const int kk = l / 16;
    V b[2][NCT];
#pragma unroll
    for (int j = 0; j < NCT; j++) b[0][j] = *(const V *)&Dt[j * 16 + l % 16][kk * KS];
#pragma unroll
    for (int gi = 0; gi < NG; gi++) {
        const int cb = gi & 1;
        if (gi + 1 < NG) {
#pragma unroll
            for (int j = 0; j < NCT; j++)
                b[cb ^ 1][j] = *(const V *)&Dt[j * 16 + l % 16][kk * KS + (gi + 1) * VK];
        }
#pragma unroll
        for (int q = 0; q < VK; q++)
#pragma unroll
            for (int j = 0; j < NCT; j++)
                acc[j] = Mf<T>::mma(a[gi * VK + q], GeL<T>::at(b[cb][j], q), acc[j]);
    }
}

// the reduction depth KS = MP / 4 is a multiple of 4 up to GF_MAXMP / 4
template <typename T, int BN>
__device__ __forceinline__ void gE_mma_ks(int KS, typename Mf<T>::acc (&acc)[BN / 16],
                                          const T (&a)[GF_MAXMP / 4], const T (*Dt)[GeL<T>::LDT],
                                          int l)
{
    static_assert(GF_MAXMP == 128, "reduction depths below cover MP <= 128");
    switch (KS) {
    case 4: gE_mma<T, BN, 4>(acc, a, Dt, l); break;
    case 8: gE_mma<T, BN, 8>(acc, a, Dt, l); break;
    case 12: gE_mma<T, BN, 12>(acc, a, Dt, l); break;
    case 16: gE_mma<T, BN, 16>(acc, a, Dt, l); break;
    case 20: gE_mma<T, BN, 20>(acc, a, Dt, l); break;
    case 24: gE_mma<T, BN, 24>(acc, a, Dt, l); break;
    case 28: gE_mma<T, BN, 28>(acc, a, Dt, l); break;
    default: gE_mma<T, BN, 32>(acc, a, Dt, l); break;
    }
}

// E[p][c] = sum_Q (P_p|Q) D[Q][c] over columns c = (k, Ja, Ib), for the La rows p of the
// alpha string Ja (P_p = pair of entry p of Ja's alpha list).  A column tile holds BN beta
// strings of one row (k, Ja), so the tile shares Ja: wave w keeps the (P|Q) fragments of
// rows 16w .. 16w+15 of Ja in registers (La <= 64: at most four row tiles, one per wave),
// reloaded when the row changes, and the row X_k[Ja] is staged in LDS (rows of up to
// GF_XROW_BYTES: 2048 strings FP64, 4096 FP32) for the beta-list gathers.  The tile of D is built in LDS from the beta
// lists (each (Q, column) receives at most one entry); columns past the end of the row
// repeat its last column (their products are not stored).
#define GF_XROW_BYTES 16384
template <typename T, int BN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gE(
    const GfHeader *__restrict__ H, const GfTask *__restrict__ tasks)
{
    constexpr int LDT = GeL<T>::LDT;
    __shared__ __align__(16) T Dt[BN][LDT];
    constexpr int XROW = GF_XROW_BYTES / (int)sizeof(T);
    __shared__ T xs[XROW];
    constexpr int NCT = BN / 16, PARTS = 256 / BN, KSMAX = GF_MAXMP / 4;
    constexpr int EMAX = (GF_MAXLIST + PARTS - 1) / PARTS;   // list entries per thread
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n = H->n;
    const long long total = H->total[GF_E];
    const int col = tid % BN, part = tid / BN;
    const long long u0 = total * blockIdx.x / gridDim.x, u1 = total * (blockIdx.x + 1) / gridDim.x;
    if (u0 >= u1) return;
    for (int idx = tid; idx < BN * LDT; idx += 256) (&Dt[0][0])[idx] = (T)0;
    // task of the current tile and the first unit past it
    int ti = find_task(tasks, n, u0, GF_E);
    long long tend = ti + 1 < n ? tasks[ti + 1].base[GF_E] : total;
    // tile u of a task: row r = (k, Ja) = k * Na + Ja, beta strings Ib0 .. Ib0 + ncol - 1
    auto geom = [](const GfTask &t, long long u, long long &r, int &Ib0, int &ncol) {
        const long long tpr = (t.g.nb + BN - 1) / BN, q = u - t.base[GF_E];
        r = q / tpr;
        Ib0 = (int)(q - r * tpr) * BN;
        ncol = (int)std::min<long long>(BN, t.g.nb - Ib0);
    };
    // beta-list entries of this thread's column for a tile (issued one tile ahead, so their
    // global loads overlap the products of the current tile); entry i is e = part + i*PARTS
    uint32_t lcur[EMAX], lnxt[EMAX];
    auto fetch = [&](long long u, uint32_t (&dst)[EMAX]) {
        if (u >= u1) return;
        const GfTask &t = tasks[u < tend ? ti : find_task(tasks, n, u, GF_E)];
        long long r;
        int Ib0, ncol;
        geom(t, u, r, Ib0, ncol);
        const long long Ib = Ib0 + std::min(col, ncol - 1);
        const auto lst = gptr(t.g.lb + Ib * t.g.Lb);
#pragma unroll
        for (int i = 0; i < EMAX; i++) {
            const int e = part + i * PARTS;
            if (e < t.g.Lb) dst[i] = lst[e];
        }
    };
    fetch(u0, lcur);
    T a[KSMAX];
    long long arow = -1;          // (task, row) of the fragments in a[] and of the row in xs
    for (long long u = u0; u < u1; u++) {
        if (u >= tend) {
            ti = find_task(tasks, n, u, GF_E);
            tend = ti + 1 < n ? tasks[ti + 1].base[GF_E] : total;
        }
        const GfTask &t = tasks[ti];
        const GemmFci &g = t.g;
        const int MP = g.MP, KS = MP / 4, Lb = g.Lb, La = g.La;
        const bool active = 16 * w < La;
        long long r;
        int Ib0, ncol;
        geom(t, u, r, Ib0, ncol);
        const long long Nb = g.nb, k = r / g.na, Ja = r - k * g.na;
        const bool staged_row = Nb <= XROW;
        const long long key = ((long long)ti << 40) + r;
        if (key != arow) {
            // A fragment of step ks for row p = 16w + l%16 of Ja: (P_p | Q) at column
            // Q = (l/16)*KS + ks; for the GF_PDIAG entry, Ja's row of sums of (kk | Q)
            const int p = 16 * w + l % 16;
            const bool ok = active && p < La;
            const int P = ok ? (int)((gptr(g.la)[Ja * La + p] >> 16) & 0xFFu) : 0;
            const T *Arow = P == GF_PDIAG ? drows<T>(g) + Ja * MP : mpad<T>(g) + (long long)P * MP;
#pragma unroll
            for (int ks = 0; ks < KSMAX; ks++)
                a[ks] = (ok && ks < KS) ? Arow[(l / 16) * KS + ks] : (T)0;
            if (staged_row) {
                const T *__restrict__ xr = xin<T>(t, (int)k) + Ja * Nb;
                for (int i = tid; i < Nb; i += 256) xs[i] = xr[i];
            }
            arow = key;
        }
        __syncthreads();   // row staged; the previous tile's entries are cleared
        {
            // all reads first (independent, one wait), then the scatter into the D tile
            T v[EMAX];
            if (staged_row) {
#pragma unroll
                for (int i = 0; i < EMAX; i++)
                    if (part + i * PARTS < Lb) v[i] = xs[lcur[i] & 0xFFFFu];
            } else {
                const auto xr = gptr(xin<T>(t, (int)k) + Ja * Nb);
#pragma unroll
                for (int i = 0; i < EMAX; i++)
                    if (part + i * PARTS < Lb) v[i] = xr[lcur[i] & 0xFFFFu];
            }
#pragma unroll
            for (int i = 0; i < EMAX; i++)
                if (part + i * PARTS < Lb)
                    Dt[col][(lcur[i] >> 16) & 0xFFu] = (lcur[i] >> 31) ? -v[i] : v[i];
        }
        __syncthreads();   // D tile complete
        fetch(u + 1, lnxt);
        if (active) {
            typename Mf<T>::acc acc[NCT];
#pragma unroll
            for (int j = 0; j < NCT; j++) acc[j] = Mf<T>::zero();
            gE_mma_ks<T, BN>(KS, acc, a, Dt, l);
            const auto E = gptr_w((T *)t.E);
            const long long c0 = r * Nb + Ib0;
#pragma unroll
            for (int j = 0; j < NCT; j++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int p = 16 * w + Mf<T>::row(l, v), cj = j * 16 + l % 16;
                    if (p < La && cj < ncol) E[(long long)p * t.lde + c0 + cj] = acc[j][v];
                }
        }
        __syncthreads();   // the products have read the D tile
        // clear exactly the entries this thread wrote (the rest of Dt stays zero)
#pragma unroll
        for (int i = 0; i < EMAX; i++)
            if (part + i * PARTS < Lb) Dt[col][(lcur[i] >> 16) & 0xFFu] = (T)0;
#pragma unroll
        for (int i = 0; i < EMAX; i++) lcur[i] = lnxt[i];
    }
}

// y_k[Ia][Ib] += sum over the alpha list of Ia of sign * E[p][k, Ja, Ib], p the entry's E row
// (bits 24..30); 256 elements per unit
template <typename T>
__global__ __launch_bounds__(256) void k_gga(const GfHeader *__restrict__ H,
                                             const GfTask *__restrict__ tasks)
{
    const int n = H->n;
    const long long total = H->total[GF_GA];
    for (long long u = blockIdx.x; u < total; u += gridDim.x) {
        const GfTask &t = tasks[find_task(tasks, n, u, GF_GA)];
        const GemmFci &g = t.g;
        const long long Nb = g.nb, N = g.na * g.nb, ncols = N * t.K, lde = t.lde;
        const long long idx = (u - t.base[GF_GA]) * 256 + threadIdx.x;
        if (idx >= ncols) continue;
        const long long k = idx / N, rem = idx - k * N;
        const long long Ia = rem / Nb, Ib = rem - Ia * Nb;
        const auto lst = gptr(g.la + Ia * g.La);
        const auto Ek = gptr((const T *)t.E + k * N + Ib);
        T acc = (T)0;
        int e = 0;
        // eight independent loads in flight per lane
        for (; e + 8 <= g.La; e += 8) {
            T v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t wd = lst[e + q];
                v[q] = Ek[(long long)((wd >> 24) & 0x7Fu) * lde + (long long)(wd & 0xFFFFu) * Nb];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) acc += (lst[e + q] >> 31) ? -v[q] : v[q];
        }
        for (; e < g.La; e++) {
            const uint32_t wd = lst[e];
            const T v = Ek[(long long)((wd >> 24) & 0x7Fu) * lde + (long long)(wd & 0xFFFFu) * Nb];
            acc += (wd >> 31) ? -v : v;
        }
        t.y[k * t.ldy + rem] += (double)acc;
    }
}

// ---- launchers ------------------------------------------------------------------------------
int jk_gf_dense(long long nstr, const uint64_t *mask, const long long *off1, const int2 *e1,
                const double *v1, const long long *off2, const int2 *e2, const double *v2,
                const double *m1, const double *m2, double *H, hipStream_t st)
{
    if (hipMemsetAsync(H, 0, sizeof(double) * (size_t)nstr * nstr, st) != hipSuccess) {
        jci_set_error("gf_dense: memset failed");
        return -2;
    }
    hipLaunchKernelGGL(k_gf_dense, dim3((unsigned)((nstr + 127) / 128)), dim3(128), 0, st, nstr,
                       mask, off1, e1, v1, off2, e2, v2, m1, m2, H);
    GF_CHECK();
    return 0;
}

int jk_gf_lists(long long nstr, const uint64_t *mask, const long long *off1, const int2 *e1,
                int L, int diag_one, uint32_t *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_gf_lists, dim3((unsigned)((nstr + 127) / 128)), dim3(128), 0, st, nstr,
                       mask, off1, e1, L, diag_one, out);
    GF_CHECK();
    return 0;
}

int jk_gf_rank(long long nstr, int L, const uint32_t *in, uint32_t *out, int *missing,
               hipStream_t st)
{
    if (nstr * L == 0) return 0;
    hipLaunchKernelGGL(k_gf_rank, dim3((unsigned)((nstr * L + 255) / 256)), dim3(256), 0, st, nstr, L,
                       in, out, missing);
    GF_CHECK();
    return 0;
}

int jk_gf_m2pad(int m1, int MP, const double *m2sq, double *o64, float *o32, hipStream_t st)
{
    hipLaunchKernelGGL(k_gf_m2pad, dim3((MP * MP + 255) / 256), dim3(256), 0, st, m1, MP, m2sq,
                       o64, o32);
    GF_CHECK();
    return 0;
}

int jk_gf_diagrows(long long nstr, int MP, const uint64_t *mask, const double *M64, double *o64,
                   float *o32, hipStream_t st)
{
    if (nstr * MP == 0) return 0;
    hipLaunchKernelGGL(k_gf_diagrows, dim3((unsigned)((nstr * MP + 255) / 256)), dim3(256), 0, st,
                       nstr, MP, mask, M64, o64, o32);
    GF_CHECK();
    return 0;
}

int jk_gf_bn_valid(int prec, int bn)
{
    return prec == 64 ? (bn == 16 || bn == 32) : (bn == 32 || bn == 64);
}

void jk_gf_units(int prec, int bn, const GfTask &t, long long units[4])
{
    const long long na = t.g.na, nb = t.g.nb, N = na * nb, cols = N * t.K;
    units[GF_CV] = prec == 64 ? 0 : (cols + 255) / 256;
    units[GF_SS] = (long long)t.K * ((na + 63) / 64) * ((nb + 63) / 64);
    units[GF_E] = (long long)t.K * na * ((nb + bn - 1) / bn);
    units[GF_GA] = (cols + 255) / 256;
}

// persistent grid caps of the grid-stride kernels (conversion, gather)
static const long long gf_cap[4] = {4096, 2048, 2048, 4096};

// blocks of 256 threads of a kernel resident on the whole device at once
static long long resident_blocks(const void *kernel)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess)
        return 0;
    return (long long)cus * per;
}

template <typename T, int BN>
static int gf_group_t(const GfHeader *d_hdr, const GfTask *d_tasks, const GfHeader *hh,
                      hipStream_t st)
{
    // The E kernel splits its units into equal consecutive ranges, one per block, and the
    // same-spin kernel walks whole 64 x 64 tiles: both run one wave of resident blocks
    // (a grid past residency leaves a last, partly filled wave of blocks)
    static const long long cap_e = std::max(256LL, resident_blocks((const void *)k_gE<T, BN>));
    static const long long cap_ss = std::max(256LL, resident_blocks((const void *)k_gss<T>));
    const long long cap[4] = {gf_cap[GF_CV], cap_ss, cap_e, gf_cap[GF_GA]};
    unsigned g[4];
    for (int i = 0; i < 4; i++)
        g[i] = (unsigned)(hh ? std::max<long long>(1, std::min(hh->total[i], cap[i])) : cap[i]);
    if (sizeof(T) == 4) {
        hipLaunchKernelGGL(k_gcv, dim3(g[GF_CV]), dim3(256), 0, st, d_hdr, d_tasks);
        GF_CHECK();
    }
    hipLaunchKernelGGL((k_gE<T, BN>), dim3(g[GF_E]), dim3(256), 0, st, d_hdr, d_tasks);
    GF_CHECK();
    hipLaunchKernelGGL(k_gss<T>, dim3(g[GF_SS]), dim3(256), 0, st, d_hdr, d_tasks);
    GF_CHECK();
    hipLaunchKernelGGL(k_gga<T>, dim3(g[GF_GA]), dim3(256), 0, st, d_hdr, d_tasks);
    GF_CHECK();
    return 0;
}

int jk_gf_group(int prec, int bn, const GfHeader *d_hdr, const GfTask *d_tasks,
                const GfHeader *hdr_host, hipStream_t st)
{
    if (prec == 64) {
        if (bn == 16) return gf_group_t<double, 16>(d_hdr, d_tasks, hdr_host, st);
        if (bn == 32) return gf_group_t<double, 32>(d_hdr, d_tasks, hdr_host, st);
    } else {
        if (bn == 32) return gf_group_t<float, 32>(d_hdr, d_tasks, hdr_host, st);
        if (bn == 64) return gf_group_t<float, 64>(d_hdr, d_tasks, hdr_host, st);
    }
    jci_set_error("gf_sigma: E tile width %d not available for FP%d", bn, prec);
    return -1;
}
