// Handle API of libjayci_gpu: device context, determinant space, Hamiltonian (integrals,
// diagonal, optional pair list), sigma dispatch and the block Davidson driver.
#include "jci_internal.h"
#include "slater.h"
#include "sigma_fast.h"
#include "sigma_gemm.h"
#include "../../include/jayci_gpu.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <algorithm>
#include <vector>

// ---- errors ---------------------------------------------------------------------------
static thread_local char g_err[1024];

void jci_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

JCI_API const char *jci_last_error(void) { return g_err; }

#define RCHECK(x)                      \
    do {                               \
        int rc_ = (x);                 \
        if (rc_) return rc_;           \
    } while (0)

static double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// ---- extra kernels used here (defined in kernels.hip) ----------------------------------
int jk_det_masks(const SpaceDesc *d_desc, const uint64_t *ma, const uint64_t *mb,
                 const long long *idx, long long first, long long count, uint64_t *oa,
                 uint64_t *ob, hipStream_t st);
int jk_rotate_inplace(int m, int kc, long long N, double *V, long long ld, const double *Y,
                      hipStream_t st);
int jk_scatter_guess(int nb, int nvec, const long long *idx, const double *Yb, double *V,
                     long long ld, hipStream_t st);
int jk_index_of(const SpaceDesc *d_desc, const unsigned long long *binom, uint64_t a,
                uint64_t b, long long *out, hipStream_t st);
int jk_copy(long long bytes, const void *src, void *dst, hipStream_t st);
int jk_fill_random(long long n, unsigned long long seed, double *x, hipStream_t st);

// ---- context ----------------------------------------------------------------------------


JCI_API int jci_device_count(int *n)
{
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *n = c;
    return 0;
}

JCI_API int jci_device_memory(int device, size_t *free_b, size_t *total_b)
{
    size_t f = 0, t = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) {
        jci_set_error("device %d: no memory information", device);
        return -2;
    }
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return 0;
}

JCI_API int jci_ctx_create(int device, jci_ctx **out)
{
    *out = nullptr;
    int n = 0;
    jci_device_count(&n);
    if (n <= 0) { jci_set_error("no HIP device visible"); return -2; }
    if (device < 0 || device >= n) { jci_set_error("device %d out of range", device); return -1; }
    JCI_CHECK(hipSetDevice(device));
    hipStream_t st = nullptr;
    JCI_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    jci_ctx *c = (jci_ctx *)calloc(1, sizeof(jci_ctx));
    c->device = device;
    c->stream = st;
    JCI_CHECK(hipMalloc(&c->d_partial, sizeof(double) * JCI_PARTIAL_BLOCKS * JCI_MAXBASIS * JCI_MAXBASIS));
    JCI_CHECK(hipMalloc(&c->d_small, sizeof(double) * JCI_MAXBASIS * JCI_MAXBASIS));
    JCI_CHECK(hipHostMalloc(&c->h_small, sizeof(double) * JCI_MAXBASIS * JCI_MAXBASIS, 0));
    std::vector<unsigned long long> tab(65 * 65);
    jci_binom_table(tab.data());
    JCI_CHECK(hipMalloc(&c->d_binom, sizeof(unsigned long long) * 65 * 65));
    JCI_CHECK(hipMemcpy(c->d_binom, tab.data(), sizeof(unsigned long long) * 65 * 65,
                        hipMemcpyHostToDevice));
    int rc = jci_noalloc_setup(device);
    if (rc) return rc;
    *out = c;
    return 0;
}

static void ctx_free(jci_ctx *c)
{
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipFree(c->d_partial);
    hipFree(c->d_small);
    hipFree(c->d_binom);
    hipHostFree(c->h_small);
    hipStreamDestroy(c->stream);
    free(c);
}

// a context is freed when it was destroyed and no space refers to it any more
JCI_API void jci_ctx_destroy(jci_ctx *c)
{
    if (!c) return;
    if (__atomic_load_n(&c->refs, __ATOMIC_ACQUIRE) > 0) { c->closing = 1; return; }
    ctx_free(c);
}

JCI_API int jci_ctx_sync(jci_ctx *c)
{
    JCI_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

JCI_API void *jci_ctx_stream(jci_ctx *c) { return (void *)c->stream; }

// ---- space --------------------------------------------------------------------------------
struct jci_space {
    jci_ctx *ctx;
    int refs, closing;          // Hamiltonians holding this space
    SpaceDesc h;
    SpaceDesc *d_desc;
    uint64_t *d_mask[2];
    int na, nb, xlvl, rule;
};

JCI_API int jci_space_create(jci_ctx *ctx, int aelec, int belec, int norbs, int ndocc,
                                int nactv, int xlvl, int rule, jci_space **out)
{
    *out = nullptr;
    jci_space *s = (jci_space *)calloc(1, sizeof(jci_space));
    s->ctx = ctx; s->na = aelec; s->nb = belec; s->xlvl = xlvl; s->rule = rule;
    if (jci_build_desc(aelec, belec, norbs, ndocc, nactv, xlvl, rule, &s->h)) { free(s); return -1; }
    if (s->h.ndets <= 0) { free(s); jci_set_error("space: no determinants"); return -1; }
    JCI_CHECK(hipSetDevice(ctx->device));
    JCI_CHECK(hipMalloc(&s->d_desc, sizeof(SpaceDesc)));
    JCI_CHECK(hipMemcpyAsync(s->d_desc, &s->h, sizeof(SpaceDesc), hipMemcpyHostToDevice, ctx->stream));
    for (int sp = 0; sp < 2; sp++) {
        long long n = s->h.s[sp].gstart[s->h.s[sp].ngrp];
        JCI_CHECK(hipMalloc(&s->d_mask[sp], sizeof(uint64_t) * (n > 0 ? n : 1)));
        RCHECK(jk_gen_strings(s->d_desc, s->h, sp, ctx->d_binom, s->d_mask[sp], ctx->stream));
    }
    JCI_CHECK(hipStreamSynchronize(ctx->stream));
    // contexts may create spaces on one host thread and free them on another
    __atomic_add_fetch(&ctx->refs, 1, __ATOMIC_ACQ_REL);
    *out = s;
    return 0;
}

static void space_free(jci_space *s)
{
    hipSetDevice(s->ctx->device);
    hipStreamSynchronize(s->ctx->stream);
    hipFree(s->d_desc);
    hipFree(s->d_mask[0]);
    hipFree(s->d_mask[1]);
    jci_ctx *c = s->ctx;
    free(s);
    if (__atomic_sub_fetch(&c->refs, 1, __ATOMIC_ACQ_REL) == 0 && c->closing) ctx_free(c);
}

// a space is freed when it was destroyed and no Hamiltonian refers to it any more
JCI_API void jci_space_destroy(jci_space *s)
{
    if (!s) return;
    if (s->refs > 0) { s->closing = 1; return; }
    space_free(s);
}

JCI_API long long jci_space_ndets(const jci_space *s) { return s->h.ndets; }
JCI_API long long jci_space_nstrings(const jci_space *s, int spin)
{
    return s->h.s[spin].gstart[s->h.s[spin].ngrp];
}
JCI_API int jci_space_ngroups(const jci_space *s, int spin) { return s->h.s[spin].ngrp; }
JCI_API int jci_space_groups(const jci_space *s, int spin, long long *out)
{
    const SpinDesc &S = s->h.s[spin];
    for (int g = 0; g < S.ngrp; g++) {
        out[5 * g] = S.gd[g]; out[5 * g + 1] = S.ga[g]; out[5 * g + 2] = S.gv[g];
        out[5 * g + 3] = S.gstart[g]; out[5 * g + 4] = S.gn[g];
    }
    return 0;
}
JCI_API int jci_space_nblocks(const jci_space *s) { return s->h.nblk; }
JCI_API int jci_space_blocks(const jci_space *s, long long *out)
{
    for (int b = 0; b < s->h.nblk; b++) {
        out[3 * b] = s->h.bp[b]; out[3 * b + 1] = s->h.bq[b]; out[3 * b + 2] = s->h.boff[b];
    }
    return 0;
}

JCI_API int jci_space_strings(jci_space *s, int spin, uint64_t *masks)
{
    JCI_CHECK(hipSetDevice(s->ctx->device));
    long long n = jci_space_nstrings(s, spin);
    JCI_CHECK(hipMemcpyAsync(masks, s->d_mask[spin], sizeof(uint64_t) * n, hipMemcpyDeviceToHost,
                             s->ctx->stream));
    JCI_CHECK(hipStreamSynchronize(s->ctx->stream));
    return 0;
}

JCI_API int jci_space_dets(jci_space *s, long long first, long long count, uint64_t *a,
                              uint64_t *b)
{
    if (first < 0 || count < 0 || first + count > s->h.ndets) {
        jci_set_error("space_dets: range out of bounds");
        return -1;
    }
    if (count == 0) return 0;
    JCI_CHECK(hipSetDevice(s->ctx->device));
    uint64_t *d;
    JCI_CHECK(hipMalloc(&d, sizeof(uint64_t) * 2 * count));
    RCHECK(jk_det_masks(s->d_desc, s->d_mask[0], s->d_mask[1], nullptr, first, count, d,
                        d + count, s->ctx->stream));
    JCI_CHECK(hipMemcpyAsync(a, d, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, s->ctx->stream));
    JCI_CHECK(hipMemcpyAsync(b, d + count, sizeof(uint64_t) * count, hipMemcpyDeviceToHost,
                             s->ctx->stream));
    JCI_CHECK(hipStreamSynchronize(s->ctx->stream));
    JCI_CHECK(hipFree(d));
    return 0;
}

JCI_API long long jci_space_index(jci_space *s, uint64_t a, uint64_t b)
{
    if (hipSetDevice(s->ctx->device) != hipSuccess) return -1;
    long long *d, r = -1;
    if (hipMalloc(&d, sizeof(long long)) != hipSuccess) return -1;
    if (jk_index_of(s->d_desc, s->ctx->d_binom, a, b, d, s->ctx->stream) == 0) {
        hipMemcpyAsync(&r, d, sizeof(long long), hipMemcpyDeviceToHost, s->ctx->stream);
        hipStreamSynchronize(s->ctx->stream);
    }
    hipFree(d);
    return r;
}

// ---- Hamiltonian --------------------------------------------------------------------------
struct jci_ham {
    jci_space *sp;
    double *d_m1, *d_m2;
    double *d_hd;
    long long *d_rowptr;
    int *d_col;
    double *d_val64;
    float *d_val32;
    long long nnz;
    int upper;
    float *d_xf;             // FP32 staging for sigma inputs
    long long xf_len;
    // excitation tables of the matrix-free sigma (spin 1 aliases spin 0 when equal)
    int have_fast;
    long long *d_off1[2], *d_off2[2];
    int *d_goff1[2], *d_goff2[2];
    int2 *d_e1[2], *d_e2[2];
    double *d_v1[2], *d_v2[2];
    long long nexc1[2], nexc2[2];
    double *d_m2sq;
    long long *d_tile_off[3];   // tile maps for W = 64, 128, 256
    FastArgs fast[3];
    int fast_w;                 // index into fast[] used by the FAST methods
    // tile-width classes: each block in one launch of the width that fits its strings per
    // row (64 .. 512), used when fast_w == 3
    FastArgs fcls[4];
    long long *d_cls_off[4];
    // the beta side as a second pass over the block-transposed vector (jk_sigma_fast_split)
    int fast_split;             // 0: one fused kernel; two passes: 1 every block in its own
                                // orientation, 2 blocks of short beta rows transposed,
                                // 3 every block transposed
    // two-pass launches in width-class mode (set_orient): main pass (one spin's side and the
    // alpha-beta products) and complement (the other spin's side) per width class, for the
    // blocks kept in their orientation [0, 4) and those run transposed [4, 8)
    FastArgs fm[8], fc[8];
    long long *d_ooff[16];
    int fast_xg;                // consecutive tiles per XCD run (jk sigma_fast xcd_tile)
    int fast_cm;                // chunk-major tile order in a block (sigma_fast tile_pos)
    SpaceDesc *d_descT;         // transposed descriptor (spins swapped)
    long long *d_tileT_off[3];
    FastArgs fastT[3];
    long long *d_tt_off, ntt;   // 16 x 16 transpose tiles per block
    void *d_xT;                 // block-transposed input
    size_t xT_bytes;
    int have_fci;               // single-block space with n <= 16: FCI kernel usable
    FciArgs fci;
    int tuned64, tuned32;       // autotuned default methods (0 = not tuned)
    // dense-string MFMA sigma (full-CI spaces)
    int have_gf;
    GemmFci gf;
    double *d_gfH64[2];
    float *d_gfH32[2];
    double *d_gfM64, *d_gfD64;
    float *d_gfM32, *d_gfD32;
    uint32_t *d_gfla, *d_gflb;
    void *d_gfE;                // opposite-spin intermediate E
    size_t gfE_bytes;
    float *d_gfX32;             // FP32 copies of the input vectors, vector-major
    size_t gfX32_len;
    double *d_gfT;              // vector-major staging for determinant-major callers
    size_t gfT_len;
    // launch descriptors of the single-instance path: one slot per vector chunk, host copies
    // pinned; a slot is uploaded only when its content changes
    GfTask *h_gft, *d_gft;
    GfHeader *h_gfh, *d_gfh;
    int gf_valid[JCI_MAXK];
    int gf_bn64, gf_bn32;       // E tile widths (autotuned)
    // initial block computed ahead (jci_ham_prepare_guess): its determinants and the lowest
    // g_nguess eigenvectors of H on them (vector j at g_U + j*g_nb), host memory
    long long *g_idx;
    double *g_U;
    int g_nb, g_nguess, g_refdim, g_diag;
};

static size_t m1_len(int n) { return (size_t)n * (n + 1) / 2; }
static size_t m2_len(int n) { size_t a = m1_len(n); return a * (a + 1) / 2; }

static int build_pairs(jci_ham *h, int flags)
{
    jci_space *s = h->sp;
    jci_ctx *c = s->ctx;
    long long N = s->h.ndets;
    if (N >= (1LL << 31) - 1) {
        jci_set_error("pair list: %lld determinants exceed 32-bit column indices", N);
        return -1;
    }
    h->upper = (flags & JCI_PAIRS_UPPER) ? 1 : 0;
    long long *d_cnt;
    JCI_CHECK(hipMalloc(&d_cnt, sizeof(long long) * N));
    JCI_CHECK(hipMalloc(&h->d_rowptr, sizeof(long long) * (N + 1)));
    RCHECK(jk_row_count(s->d_desc, N, s->d_mask[0], s->d_mask[1], c->d_binom, d_cnt, h->upper,
                        c->stream));
    RCHECK(jk_scan(d_cnt, h->d_rowptr, N, c->stream));
    JCI_CHECK(hipMemcpyAsync(&h->nnz, h->d_rowptr + N, sizeof(long long), hipMemcpyDeviceToHost,
                             c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    JCI_CHECK(hipFree(d_cnt));
    size_t nz = (size_t)(h->nnz > 0 ? h->nnz : 1);
    if (hipMalloc(&h->d_col, sizeof(int) * nz) != hipSuccess) {
        jci_set_error("pair list: out of device memory (%lld pairs)", h->nnz);
        return -3;
    }
    if ((flags & JCI_PAIRS_FP64) && hipMalloc(&h->d_val64, sizeof(double) * nz) != hipSuccess) {
        jci_set_error("pair list: out of device memory (%lld FP64 values)", h->nnz);
        return -3;
    }
    if ((flags & JCI_PAIRS_FP32) && hipMalloc(&h->d_val32, sizeof(float) * nz) != hipSuccess) {
        jci_set_error("pair list: out of device memory (%lld FP32 values)", h->nnz);
        return -3;
    }
    RCHECK(jk_row_fill(s->d_desc, N, s->d_mask[0], s->d_mask[1], c->d_binom, h->d_m1, h->d_m2,
                       h->d_rowptr, h->d_col, h->d_val64, h->d_val32, h->upper, c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// dense same-spin matrices, replacement lists and padded (P|Q) of the dense-string sigma;
// taken when the space is one alpha group times one beta group (full CI), n <= 15 and the
// dense matrices stay below 8 GB
static int build_gf(jci_ham *h)
{
    jci_space *s = h->sp;
    jci_ctx *c = s->ctx;
    const SpaceDesc &D = s->h;
    if (D.nblk != 1 || D.s[0].ngrp != 1 || D.s[1].ngrp != 1) return 0;
    if (D.s[0].gstart[0] != 0 || D.s[1].gstart[0] != 0) return 0;
    const int n = D.norbs, m1 = n * (n + 1) / 2, MP = (m1 + 15) / 16 * 16;
    const long long na = D.s[0].gn[0], nb = D.s[1].gn[0];
    if (MP > GF_MAXMP || na > GF_MAXSTR || nb > GF_MAXSTR) return 0;
    if (D.allmask != (n >= 64 ? ~0ULL : ((1ULL << n) - 1ULL))) return 0;
    const bool same = s->na == s->nb;
    if ((double)(na * na + (same ? 0 : nb * nb)) * 12.0 > 8e9) return 0;
    const int sa = s->na * (n - s->na), sb = s->nb * (n - s->nb);
    if (h->nexc1[0] != na * sa || h->nexc1[1] != nb * sb) return 0;
    // alpha lists: the singles and one entry for the occupied diagonal pairs together
    const int La = sa + (s->na > 0 ? 1 : 0), Lb = sb + s->nb;
    if (La > GF_MAXLIST || Lb > GF_MAXLIST) return 0;
    JCI_CHECK(hipMalloc(&h->d_gfH64[0], sizeof(double) * na * na));
    JCI_CHECK(hipMalloc(&h->d_gfH32[0], sizeof(float) * na * na));
    RCHECK(jk_gf_dense(na, s->d_mask[0], h->d_off1[0], h->d_e1[0], h->d_v1[0], h->d_off2[0],
                       h->d_e2[0], h->d_v2[0], h->d_m1, h->d_m2, h->d_gfH64[0], c->stream));
    RCHECK(jk_to_f32(h->d_gfH64[0], h->d_gfH32[0], na * na, c->stream));
    if (same) {
        h->d_gfH64[1] = h->d_gfH64[0];
        h->d_gfH32[1] = h->d_gfH32[0];
    } else {
        JCI_CHECK(hipMalloc(&h->d_gfH64[1], sizeof(double) * nb * nb));
        JCI_CHECK(hipMalloc(&h->d_gfH32[1], sizeof(float) * nb * nb));
        RCHECK(jk_gf_dense(nb, s->d_mask[1], h->d_off1[1], h->d_e1[1], h->d_v1[1], h->d_off2[1],
                           h->d_e2[1], h->d_v2[1], h->d_m1, h->d_m2, h->d_gfH64[1], c->stream));
        RCHECK(jk_to_f32(h->d_gfH64[1], h->d_gfH32[1], nb * nb, c->stream));
    }
    JCI_CHECK(hipMalloc(&h->d_gfM64, sizeof(double) * MP * MP));
    JCI_CHECK(hipMalloc(&h->d_gfM32, sizeof(float) * MP * MP));
    RCHECK(jk_gf_m2pad(m1, MP, h->d_m2sq, h->d_gfM64, h->d_gfM32, c->stream));
    JCI_CHECK(hipMalloc(&h->d_gfD64, sizeof(double) * na * MP));
    JCI_CHECK(hipMalloc(&h->d_gfD32, sizeof(float) * na * MP));
    RCHECK(jk_gf_diagrows(na, MP, s->d_mask[0], h->d_gfM64, h->d_gfD64, h->d_gfD32, c->stream));
    JCI_CHECK(hipMalloc(&h->d_gfla, sizeof(uint32_t) * na * La));
    JCI_CHECK(hipMalloc(&h->d_gflb, sizeof(uint32_t) * nb * Lb));
    RCHECK(jk_gf_lists(nb, s->d_mask[1], h->d_off1[1], h->d_e1[1], Lb, 0, h->d_gflb, c->stream));
    {
        // alpha lists with the E row of each entry (the pair's position in the target's list)
        uint32_t *raw = nullptr;
        int *d_missing = nullptr, missing = 0;
        // temporaries stream-ordered: a hipFree would wait for every stream of the device
        JCI_CHECK(hipMallocAsync((void **)&raw, sizeof(uint32_t) * na * La, c->stream));
        JCI_CHECK(hipMallocAsync((void **)&d_missing, sizeof(int), c->stream));
        JCI_CHECK(hipMemsetAsync(d_missing, 0, sizeof(int), c->stream));
        int rc = jk_gf_lists(na, s->d_mask[0], h->d_off1[0], h->d_e1[0], La, 1, raw, c->stream);
        if (!rc) rc = jk_gf_rank(na, La, raw, h->d_gfla, d_missing, c->stream);
        if (!rc && hipMemcpyAsync(&missing, d_missing, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = -2;
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = -2;
        hipFreeAsync(raw, c->stream);
        hipFreeAsync(d_missing, c->stream);
        if (rc) return rc;
        if (missing) {
            jci_set_error("gf: %d alpha-list entries without a row in their target's list", missing);
            return -1;
        }
    }
    JCI_CHECK(hipStreamSynchronize(c->stream));
    GemmFci &g = h->gf;
    g.na = na;
    g.nb = nb;
    g.m1 = m1;
    g.MP = MP;
    g.La = La;
    g.Lb = Lb;
    g.H64[0] = h->d_gfH64[0];
    g.H64[1] = h->d_gfH64[1];
    g.H32[0] = h->d_gfH32[0];
    g.H32[1] = h->d_gfH32[1];
    g.M64 = h->d_gfM64;
    g.M32 = h->d_gfM32;
    g.la = h->d_gfla;
    g.lb = h->d_gflb;
    g.D64 = h->d_gfD64;
    g.D32 = h->d_gfD32;
    h->gf_bn64 = GF_BN64_DEFAULT;
    h->gf_bn32 = GF_BN32_DEFAULT;
    h->have_gf = 1;
    return 0;
}

// excitation tables + tile map of the matrix-free sigma
// Two-pass launch lists of the width-class mode for the orientation h->fast_split asks for.
// A block of na alpha x nq beta strings runs in its own orientation (main pass: one alpha
// string per tile, lanes over beta strings, on x; complement: the beta side on the
// transposed descriptor over xT) or transposed (main pass on the transposed descriptor over
// xT, lanes over alpha strings; complement: the alpha side on the original descriptor over
// x).  Mode 2 transposes the blocks whose beta rows are short (nq < 256) and whose alpha
// columns are longer, so that their tiles fill; every launch is sized by its lane count.
static int set_orient(jci_ham *h)
{
    jci_space *s = h->sp;
    jci_ctx *c = s->ctx;
    const SpaceDesc &D = s->h;
    const int mode = h->fast_split;
    const int cw[4] = {64, 128, 256, 512};
    auto cls_of = [&](long long n) { return n <= 64 ? 0 : n <= 128 ? 1 : n <= 256 ? 2 : 3; };
    auto tiles = [&](long long rows, long long lanes, int ci) {
        return cls_of(lanes) == ci ? rows * ((lanes + cw[ci] - 1) / cw[ci]) : 0LL;
    };
    std::vector<long long> off(16 * (size_t)(D.nblk + 1), 0);
    for (int b = 0; b < D.nblk; b++) {
        const long long na = D.s[0].gn[D.bp[b]], nq = D.s[1].gn[D.bq[b]];
        const bool tr = mode == 3 || (mode == 2 && nq < 256 && na > nq);
        for (int ci = 0; ci < 4; ci++) {
            const long long t[4] = {tr ? 0 : tiles(na, nq, ci),    // main, own orientation
                                    tr ? tiles(nq, na, ci) : 0,    // main, transposed
                                    tr ? 0 : tiles(nq, na, ci),    // complement of the first
                                    tr ? tiles(na, nq, ci) : 0};   // complement of the second
            for (int k = 0; k < 4; k++) {
                long long *o = &off[(size_t)(k * 4 + ci) * (D.nblk + 1)];
                o[b + 1] = o[b] + t[k];
            }
        }
    }
    for (int i = 0; i < 16; i++)
        JCI_CHECK(hipMemcpyAsync(h->d_ooff[i], &off[(size_t)i * (D.nblk + 1)],
                                 sizeof(long long) * (D.nblk + 1), hipMemcpyHostToDevice,
                                 c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    for (int ci = 0; ci < 4; ci++) {
        const FastArgs *base[4] = {&h->fast[2], &h->fastT[2], &h->fastT[2], &h->fast[2]};
        FastArgs *dst[4] = {&h->fm[ci], &h->fm[4 + ci], &h->fc[ci], &h->fc[4 + ci]};
        for (int k = 0; k < 4; k++) {
            FastArgs &a = *dst[k];
            a = *base[k];
            a.W = cw[ci];
            a.tile_off = h->d_ooff[k * 4 + ci];
            a.ntiles = off[(size_t)(k * 4 + ci) * (D.nblk + 1) + D.nblk];
            a.tout = (k == 1 || k == 2);
            a.src_t = (k == 1 || k == 2);
            a.hd = h->d_hd;
        }
    }
    return 0;
}

static int build_fast(jci_ham *h, int W)
{
    jci_space *s = h->sp;
    jci_ctx *c = s->ctx;
    const SpaceDesc &D = s->h;
    bool same = s->na == s->nb;
    for (int spin = 0; spin < (same ? 1 : 2); spin++) {
        long long n = D.s[spin].gstart[D.s[spin].ngrp];
        long long *c1, *c2;
        int G = D.s[spin].ngrp;
        JCI_CHECK(hipMallocAsync((void **)&c1, sizeof(long long) * (n + 1), c->stream));
        JCI_CHECK(hipMallocAsync((void **)&c2, sizeof(long long) * (n + 1), c->stream));
        JCI_CHECK(hipMalloc(&h->d_off1[spin], sizeof(long long) * (n + 1)));
        JCI_CHECK(hipMalloc(&h->d_off2[spin], sizeof(long long) * (n + 1)));
        JCI_CHECK(hipMalloc(&h->d_goff1[spin], sizeof(int) * n * (G + 1)));
        JCI_CHECK(hipMalloc(&h->d_goff2[spin], sizeof(int) * n * (G + 1)));
        RCHECK(jk_exc_count(s->d_desc, spin, n, s->d_mask[spin], c->d_binom, c1, c2,
                            h->d_goff1[spin], h->d_goff2[spin], c->stream));
        RCHECK(jk_scan(c1, h->d_off1[spin], n, c->stream));
        RCHECK(jk_scan(c2, h->d_off2[spin], n, c->stream));
        JCI_CHECK(hipMemcpyAsync(&h->nexc1[spin], h->d_off1[spin] + n, sizeof(long long),
                                 hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipMemcpyAsync(&h->nexc2[spin], h->d_off2[spin] + n, sizeof(long long),
                                 hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
        hipFreeAsync(c1, c->stream);
        hipFreeAsync(c2, c->stream);
        size_t a1 = h->nexc1[spin] > 0 ? h->nexc1[spin] : 1, a2 = h->nexc2[spin] > 0 ? h->nexc2[spin] : 1;
        if (hipMalloc(&h->d_e1[spin], sizeof(int2) * a1) != hipSuccess ||
            hipMalloc(&h->d_v1[spin], sizeof(double) * a1) != hipSuccess ||
            hipMalloc(&h->d_e2[spin], sizeof(int2) * a2) != hipSuccess ||
            hipMalloc(&h->d_v2[spin], sizeof(double) * a2) != hipSuccess) {
            jci_set_error("excitation tables: out of device memory (%lld + %lld entries)",
                          h->nexc1[spin], h->nexc2[spin]);
            return -3;
        }
        RCHECK(jk_exc_fill(s->d_desc, spin, n, s->d_mask[spin], c->d_binom, h->d_m1, h->d_m2,
                           h->d_off1[spin], h->d_off2[spin], h->d_goff1[spin], h->d_goff2[spin],
                           h->d_e1[spin], h->d_v1[spin], h->d_e2[spin], h->d_v2[spin], c->stream));
    }
    if (same) {
        h->d_off1[1] = h->d_off1[0]; h->d_off2[1] = h->d_off2[0];
        h->d_goff1[1] = h->d_goff1[0]; h->d_goff2[1] = h->d_goff2[0];
        h->d_e1[1] = h->d_e1[0]; h->d_e2[1] = h->d_e2[0];
        h->d_v1[1] = h->d_v1[0]; h->d_v2[1] = h->d_v2[0];
        h->nexc1[1] = h->nexc1[0]; h->nexc2[1] = h->nexc2[0];
    }
    int n = D.norbs, mp = n * (n + 1) / 2;
    JCI_CHECK(hipMalloc(&h->d_m2sq, sizeof(double) * (size_t)mp * mp));
    RCHECK(jk_m2_square(mp, h->d_m2, h->d_m2sq, c->stream));
    const int widths[3] = {64, 128, 256};
    for (int wi = 0; wi < 3; wi++) {
        int Wd = widths[wi];
        std::vector<long long> toff(D.nblk + 1, 0);
        for (int b = 0; b < D.nblk; b++) {
            long long na = D.s[0].gn[D.bp[b]], nq = D.s[1].gn[D.bq[b]];
            toff[b + 1] = toff[b] + na * ((nq + Wd - 1) / Wd);
        }
        JCI_CHECK(hipMalloc(&h->d_tile_off[wi], sizeof(long long) * (D.nblk + 1)));
        JCI_CHECK(hipMemcpyAsync(h->d_tile_off[wi], toff.data(), sizeof(long long) * (D.nblk + 1),
                                 hipMemcpyHostToDevice, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
        FastArgs &a = h->fast[wi];
        a.d = s->d_desc;
        a.ma = s->d_mask[0];
        a.mb = s->d_mask[1];
        a.ta = ExcView{h->d_off1[0], h->d_off2[0], h->d_e1[0], h->d_e2[0], h->d_v1[0], h->d_v2[0],
                       h->d_goff1[0], h->d_goff2[0], D.s[0].ngrp};
        a.tb = ExcView{h->d_off1[1], h->d_off2[1], h->d_e1[1], h->d_e2[1], h->d_v1[1], h->d_v2[1],
                       h->d_goff1[1], h->d_goff2[1], D.s[1].ngrp};
        a.m2sq = h->d_m2sq;
        a.hd = h->d_hd;
        a.tile_off = h->d_tile_off[wi];
        a.m1 = mp;
        a.m1int = (D.ndocc + D.nactv) * (D.ndocc + D.nactv + 1) / 2;
        a.W = Wd;
        a.A = 0;
        a.ldy = 0;
        a.xg = 0;
        a.tout = 0;
        a.src_t = 0;
        a.ntiles = toff[D.nblk];
        if (Wd == W) h->fast_w = wi;
    }
    // transposed space for the beta-side pass
    {
        SpaceDesc T;
        jci_transpose_desc(D, &T);
        JCI_CHECK(hipMalloc(&h->d_descT, sizeof(SpaceDesc)));
        JCI_CHECK(hipMemcpyAsync(h->d_descT, &T, sizeof(SpaceDesc), hipMemcpyHostToDevice, c->stream));
        std::vector<long long> tt(D.nblk + 1, 0);
        for (int b = 0; b < D.nblk; b++) {
            long long na = D.s[0].gn[D.bp[b]], nq = D.s[1].gn[D.bq[b]];
            tt[b + 1] = tt[b] + ((na + 15) / 16) * ((nq + 15) / 16);
        }
        h->ntt = tt[D.nblk];
        JCI_CHECK(hipMalloc(&h->d_tt_off, sizeof(long long) * (D.nblk + 1)));
        JCI_CHECK(hipMemcpyAsync(h->d_tt_off, tt.data(), sizeof(long long) * (D.nblk + 1),
                                 hipMemcpyHostToDevice, c->stream));
        for (int wi = 0; wi < 3; wi++) {
            const int Wd = widths[wi];
            std::vector<long long> toff(D.nblk + 1, 0);
            for (int b = 0; b < D.nblk; b++) {
                long long nbq = D.s[1].gn[D.bq[b]], na = D.s[0].gn[D.bp[b]];
                toff[b + 1] = toff[b] + nbq * ((na + Wd - 1) / Wd);
            }
            JCI_CHECK(hipMalloc(&h->d_tileT_off[wi], sizeof(long long) * (D.nblk + 1)));
            JCI_CHECK(hipMemcpyAsync(h->d_tileT_off[wi], toff.data(), sizeof(long long) * (D.nblk + 1),
                                     hipMemcpyHostToDevice, c->stream));
            JCI_CHECK(hipStreamSynchronize(c->stream));
            FastArgs &t = h->fastT[wi];
            t = h->fast[wi];
            t.d = h->d_descT;
            t.ma = s->d_mask[1];
            t.mb = s->d_mask[0];
            t.ta = h->fast[wi].tb;
            t.tb = h->fast[wi].ta;
            t.hd = nullptr;
            t.tile_off = h->d_tileT_off[wi];
            t.ntiles = toff[D.nblk];
            t.tout = 1;
            t.src_t = 1;
        }
        h->fast_split = D.nblk > 1 ? 2 : 0;
        // tiles chunk-major, in runs of 64 per XCD (16e/60o FP64 pass: 9.8 s string-major,
        // 7.8 s chunk-major, 7.5 s chunk-major in runs of 64)
        h->fast_cm = 1;
        h->fast_xg = 64;
        // spaces of several blocks: per-block widths (set below) by default
        const bool multi = D.nblk > 1;
        // width classes: rows of nq strings go to the launch of the smallest width >= nq
        // (512 above 256)
        const int cw[4] = {64, 128, 256, 512};
        auto cls_of = [&](long long nq) { return nq <= 64 ? 0 : nq <= 128 ? 1 : nq <= 256 ? 2 : 3; };
        // (the fused kernel's launches; the two-pass launches come from set_orient)
        for (int ci = 0; ci < 4; ci++) {
            std::vector<long long> toff(D.nblk + 1, 0);
            for (int b = 0; b < D.nblk; b++) {
                long long na = D.s[0].gn[D.bp[b]], nq = D.s[1].gn[D.bq[b]];
                toff[b + 1] = toff[b] + (cls_of(nq) == ci ? na * ((nq + cw[ci] - 1) / cw[ci]) : 0);
            }
            JCI_CHECK(hipMalloc(&h->d_cls_off[ci], sizeof(long long) * (D.nblk + 1)));
            JCI_CHECK(hipMemcpyAsync(h->d_cls_off[ci], toff.data(), sizeof(long long) * (D.nblk + 1),
                                     hipMemcpyHostToDevice, c->stream));
            JCI_CHECK(hipStreamSynchronize(c->stream));
            h->fcls[ci] = h->fast[2];
            h->fcls[ci].W = cw[ci];
            h->fcls[ci].tile_off = h->d_cls_off[ci];
            h->fcls[ci].ntiles = toff[D.nblk];
        }
        if (multi) h->fast_w = 3;
        for (int i = 0; i < 16; i++)
            JCI_CHECK(hipMalloc(&h->d_ooff[i], sizeof(long long) * (D.nblk + 1)));
        RCHECK(set_orient(h));
    }
    h->have_fast = 1;
    if (D.nblk == 1 && D.norbs <= FCI_MAXORB) {
        FciArgs &f = h->fci;
        f.ma = s->d_mask[0];
        f.mb = s->d_mask[1];
        f.ta = h->fast[0].ta;
        f.tb = h->fast[0].tb;
        f.m2 = h->d_m2;
        f.hd = h->d_hd;
        f.na = D.s[0].gn[D.bp[0]];
        f.nb = D.s[1].gn[D.bq[0]];
        f.m1 = mp;
        f.m2len = mp * (mp + 1) / 2;
        f.m2len_pad = (f.m2len + 3) / 4 * 4;
        long long maxs = (long long)std::max(s->na, s->nb) * (D.norbs - std::min(s->na, s->nb));
        h->have_fci = (D.s[0].gstart[0] == 0 && D.s[1].gstart[0] == 0 && maxs <= FCI_MAXSINGLES);
    }
    return build_gf(h);
}

JCI_API int jci_ham_create(jci_space *s, const double *m1, const double *m2, int flags,
                              jci_ham **out)
{
    *out = nullptr;
    jci_ctx *c = s->ctx;
    JCI_CHECK(hipSetDevice(c->device));
    jci_ham *h = (jci_ham *)calloc(1, sizeof(jci_ham));
    h->sp = s;
    s->refs++;
    int n = s->h.norbs;
    JCI_CHECK(hipMalloc(&h->d_m1, sizeof(double) * m1_len(n)));
    JCI_CHECK(hipMalloc(&h->d_m2, sizeof(double) * m2_len(n)));
    JCI_CHECK(hipMemcpyAsync(h->d_m1, m1, sizeof(double) * m1_len(n), hipMemcpyHostToDevice, c->stream));
    JCI_CHECK(hipMemcpyAsync(h->d_m2, m2, sizeof(double) * m2_len(n), hipMemcpyHostToDevice, c->stream));
    JCI_CHECK(hipMalloc(&h->d_hd, sizeof(double) * s->h.ndets));
    RCHECK(jk_hdiag(s->d_desc, s->h.ndets, s->d_mask[0], s->d_mask[1], h->d_m1, h->d_m2, h->d_hd,
                    c->stream));
    if (flags & (JCI_PAIRS_FP64 | JCI_PAIRS_FP32)) {
        int rc = build_pairs(h, flags);
        if (rc) { jci_ham_destroy(h); return rc; }
    }
    if (flags & JCI_DIRECT_TABLES) {
        int rc = build_fast(h, 256);
        if (rc) { jci_ham_destroy(h); return rc; }
    }
    JCI_CHECK(hipStreamSynchronize(c->stream));
    *out = h;
    return 0;
}

JCI_API void jci_ham_destroy(jci_ham *h)
{
    if (!h) return;
    hipSetDevice(h->sp->ctx->device);
    hipStreamSynchronize(h->sp->ctx->stream);
    hipFree(h->d_m1); hipFree(h->d_m2); hipFree(h->d_hd);
    hipFree(h->d_rowptr); hipFree(h->d_col); hipFree(h->d_val64); hipFree(h->d_val32);
    hipFree(h->d_xf);
    for (int sp = 0; sp < 2; sp++) {
        if (sp == 1 && h->d_off1[1] == h->d_off1[0]) break;
        hipFree(h->d_off1[sp]); hipFree(h->d_off2[sp]); hipFree(h->d_e1[sp]);
        hipFree(h->d_goff1[sp]); hipFree(h->d_goff2[sp]);
        hipFree(h->d_e2[sp]); hipFree(h->d_v1[sp]); hipFree(h->d_v2[sp]);
    }
    hipFree(h->d_m2sq);
    for (int wi = 0; wi < 3; wi++) hipFree(h->d_tile_off[wi]);
    for (int wi = 0; wi < 3; wi++) hipFree(h->d_tileT_off[wi]);
    for (int ci = 0; ci < 4; ci++) hipFree(h->d_cls_off[ci]);
    for (int i = 0; i < 16; i++) hipFree(h->d_ooff[i]);
    hipFree(h->d_descT);
    hipFree(h->d_tt_off);
    hipFree(h->d_xT);
    if (h->d_gfH64[1] != h->d_gfH64[0]) { hipFree(h->d_gfH64[1]); hipFree(h->d_gfH32[1]); }
    hipFree(h->d_gfH64[0]); hipFree(h->d_gfH32[0]);
    hipFree(h->d_gfM64); hipFree(h->d_gfM32); hipFree(h->d_gfla); hipFree(h->d_gflb);
    hipFree(h->d_gfD64); hipFree(h->d_gfD32);
    hipFree(h->d_gfE); hipFree(h->d_gfX32); hipFree(h->d_gfT);
    hipFree(h->d_gft); hipFree(h->d_gfh);
    if (h->h_gft) hipHostFree(h->h_gft);
    if (h->h_gfh) hipHostFree(h->h_gfh);
    free(h->g_idx);
    free(h->g_U);
    jci_space *s = h->sp;
    free(h);
    if (--s->refs == 0 && s->closing) space_free(s);
}

JCI_API long long jci_ham_npairs(const jci_ham *h) { return h->nnz; }

JCI_API long long jci_ham_string_excitations(jci_ham *h, int spin, long long string, int kind,
                                             long long cap, long long *targets, int *info,
                                             double *vals)
{
    if (!h || !h->have_fast) { jci_set_error("string_excitations: no excitation tables"); return -1; }
    if ((spin != 0 && spin != 1) || (kind != 1 && kind != 2)) {
        jci_set_error("string_excitations: spin 0/1 and kind 1/2");
        return -1;
    }
    const SpaceDesc &D = h->sp->h;
    const long long n = D.s[spin].gstart[D.s[spin].ngrp];
    if (string < 0 || string >= n) { jci_set_error("string_excitations: string out of range"); return -1; }
    jci_ctx *c = h->sp->ctx;
    JCI_CHECK(hipSetDevice(c->device));
    const long long *off = kind == 1 ? h->d_off1[spin] : h->d_off2[spin];
    long long b[2];
    JCI_CHECK(hipMemcpyAsync(b, off + string, sizeof(b), hipMemcpyDeviceToHost, c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    const long long cnt = b[1] - b[0], m = std::min(cnt, cap);
    if (m > 0 && (targets || info || vals)) {
        std::vector<int2> e(m);
        std::vector<double> v(m);
        JCI_CHECK(hipMemcpyAsync(e.data(), (kind == 1 ? h->d_e1[spin] : h->d_e2[spin]) + b[0],
                                 sizeof(int2) * m, hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipMemcpyAsync(v.data(), (kind == 1 ? h->d_v1[spin] : h->d_v2[spin]) + b[0],
                                 sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
        for (long long t = 0; t < m; t++) {
            if (targets) targets[t] = e[t].x;
            if (info) info[t] = e[t].y;
            if (vals) vals[t] = v[t];
        }
    }
    return cnt;
}

JCI_API int jci_ham_hdiag(jci_ham *h, double *out)
{
    jci_ctx *c = h->sp->ctx;
    JCI_CHECK(hipSetDevice(c->device));
    JCI_CHECK(hipMemcpyAsync(out, h->d_hd, sizeof(double) * h->sp->h.ndets, hipMemcpyDeviceToHost,
                             c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

JCI_API int jci_ham_rows(jci_ham *h, const long long *rows, int nrows, long long cap,
                            long long *cols, double *vals, int *phase, int *deg, long long *counts)
{
    jci_space *s = h->sp;
    jci_ctx *c = s->ctx;
    if (nrows <= 0) return 0;
    for (int t = 0; t < nrows; t++)
        if (rows[t] < 0 || rows[t] >= s->h.ndets) { jci_set_error("ham_rows: row out of range"); return -1; }
    JCI_CHECK(hipSetDevice(c->device));
    size_t tot = (size_t)nrows * cap;
    long long *d_rows, *d_cols, *d_cnt;
    double *d_vals;
    int *d_ph, *d_dg;
    JCI_CHECK(hipMalloc(&d_rows, sizeof(long long) * nrows));
    JCI_CHECK(hipMalloc(&d_cols, sizeof(long long) * tot));
    JCI_CHECK(hipMalloc(&d_vals, sizeof(double) * tot));
    JCI_CHECK(hipMalloc(&d_ph, sizeof(int) * tot));
    JCI_CHECK(hipMalloc(&d_dg, sizeof(int) * tot));
    JCI_CHECK(hipMalloc(&d_cnt, sizeof(long long) * nrows));
    JCI_CHECK(hipMemcpyAsync(d_rows, rows, sizeof(long long) * nrows, hipMemcpyHostToDevice, c->stream));
    RCHECK(jk_rows_debug(s->d_desc, s->d_mask[0], s->d_mask[1], c->d_binom, h->d_m1, h->d_m2,
                         d_rows, nrows, cap, d_cols, d_vals, d_ph, d_dg, d_cnt, c->stream));
    JCI_CHECK(hipMemcpyAsync(cols, d_cols, sizeof(long long) * tot, hipMemcpyDeviceToHost, c->stream));
    JCI_CHECK(hipMemcpyAsync(vals, d_vals, sizeof(double) * tot, hipMemcpyDeviceToHost, c->stream));
    JCI_CHECK(hipMemcpyAsync(phase, d_ph, sizeof(int) * tot, hipMemcpyDeviceToHost, c->stream));
    JCI_CHECK(hipMemcpyAsync(deg, d_dg, sizeof(int) * tot, hipMemcpyDeviceToHost, c->stream));
    JCI_CHECK(hipMemcpyAsync(counts, d_cnt, sizeof(long long) * nrows, hipMemcpyDeviceToHost, c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    hipFree(d_rows); hipFree(d_cols); hipFree(d_vals); hipFree(d_ph); hipFree(d_dg); hipFree(d_cnt);
    return 0;
}

static int resolve_method(const jci_ham *h, int method, int want32)
{
    if (method != JCI_SIGMA_AUTO) return method;
    if (want32 && h->tuned32) return h->tuned32;
    if (!want32 && h->tuned64) return h->tuned64;
    if (want32) {
        if (h->d_val32) return JCI_SIGMA_PAIRS32;
        if (h->have_gf) return JCI_SIGMA_GEMM32;
        if (h->have_fci) return JCI_SIGMA_FCI32;
        return h->have_fast ? JCI_SIGMA_FAST32 : JCI_SIGMA_DIRECT32;
    }
    if (h->d_val64) return h->upper ? JCI_SIGMA_PAIRS_SYM64 : JCI_SIGMA_PAIRS64;
    if (h->have_gf) return JCI_SIGMA_GEMM64;
    if (h->have_fci) return JCI_SIGMA_FCI64;
    return h->have_fast ? JCI_SIGMA_FAST64 : JCI_SIGMA_DIRECT64;
}

// a caller-provided device block carved into the buffers one sigma call needs (the
// allocation-free entry point); every carve is 256-byte aligned
int jci_scratch_carve(SigmaScratch *s, size_t bytes, void **out, const char *what)
{
    const size_t at = (s->used + 255) / 256 * 256;
    if (at > s->size || bytes > s->size - at) {
        jci_set_error("sigma: %s needs %zu bytes at offset %zu of the %zu-byte scratch array "
                      "of the allocation-free entry point", what, bytes, at, s->size);
        return -1;
    }
    *out = s->base + at;
    s->used = at + bytes;
    return 0;
}

static int ensure_xf(jci_ham *h, long long n, SigmaScratch *s, float **xf)
{
    if (s) return jci_scratch_carve(s, sizeof(float) * (size_t)n, (void **)xf, "FP32 staging");
    *xf = h->d_xf;
    if (h->xf_len >= n) return 0;
    hipFree(h->d_xf);
    h->d_xf = nullptr;
    h->xf_len = 0;
    if (hipMalloc(&h->d_xf, sizeof(float) * n) != hipSuccess) {
        jci_set_error("sigma: out of device memory for FP32 staging");
        return -3;
    }
    h->xf_len = n;
    *xf = h->d_xf;
    return 0;
}

static int grow(void **p, size_t *have, size_t want, const char *what)
{
    if (*have >= want) return 0;
    hipFree(*p);
    *p = nullptr;
    *have = 0;
    if (hipMalloc(p, want) != hipSuccess) {
        jci_set_error("sigma: out of device memory for %s (%zu bytes)", what, want);
        return -3;
    }
    *have = want;
    return 0;
}

// dense-string sigma of K vector-major vectors: y + k*ldy = H (x + k*ldx)
static int gf_run(jci_ham *h, int prec, int K, const double *x, long long ldx, double *y,
                  long long ldy, SigmaScratch *s)
{
    jci_ctx *c = h->sp->ctx;
    const long long N = h->sp->h.ndets;
    const size_t esz = prec == 64 ? 8 : 4;
    if (s) {
        // allocation-free: descriptors, the intermediate and the FP32 copies in the scratch
        // block, the vectors in passes of as many as the block holds
        GfTask *d_t;
        GfHeader *d_h;
        RCHECK(jci_scratch_carve(s, sizeof(GfTask), (void **)&d_t, "the task descriptor"));
        RCHECK(jci_scratch_carve(s, sizeof(GfHeader), (void **)&d_h, "the task header"));
        // vectors per pass: the most whose intermediate and FP32 copies fit what is left
        const size_t at = (s->used + 255) / 256 * 256;
        const size_t left = at < s->size ? s->size - at : 0;
        auto pass_bytes = [&](int k) {
            return (size_t)h->gf.La * jk_gf_lde((long long)k * N) * esz + 256 +
                   (prec == 32 ? sizeof(float) * (size_t)k * N + 256 : 0);
        };
        int Kc = K;
        while (Kc > 0 && pass_bytes(Kc) > left) Kc--;
        if (Kc < 1) {
            jci_set_error("sigma: the dense-string intermediate of %lld determinants needs %zu "
                          "bytes, more than the %zu-byte scratch array of the allocation-free "
                          "entry point holds", N, pass_bytes(1), s->size);
            return -1;
        }
        const size_t mark = s->used;
        for (int b = 0; b < K; b += Kc) {
            const int k = std::min(Kc, K - b);
            s->used = mark;
            void *E;
            float *x32 = nullptr;
            RCHECK(jci_scratch_carve(s, (size_t)h->gf.La * jk_gf_lde((long long)k * N) * esz, &E,
                                     "the dense-string intermediate"));
            if (prec == 32)
                RCHECK(jci_scratch_carve(s, sizeof(float) * (size_t)k * N, (void **)&x32,
                                         "FP32 staging"));
            GfTask t;
            GfHeader hh;
            memset(&t, 0, sizeof(t));
            memset(&hh, 0, sizeof(hh));
            t.g = h->gf;
            t.K = k;
            t.x64 = x + (size_t)b * ldx;
            t.ldx64 = ldx;
            t.x32 = x32;
            t.y = y + (size_t)b * ldy;
            t.ldy = ldy;
            t.E = E;
            t.lde = jk_gf_lde((long long)k * N);
            hh.n = 1;
            const int bn = prec == 64 ? h->gf_bn64 : h->gf_bn32;
            jk_gf_units(prec, bn, t, hh.total);
            // the descriptors come from this stack frame: wait for the copies before the
            // frame is reused (earlier passes' kernels are ordered after their own copies)
            JCI_CHECK(hipMemcpyAsync(d_t, &t, sizeof(t), hipMemcpyHostToDevice, c->stream));
            JCI_CHECK(hipMemcpyAsync(d_h, &hh, sizeof(hh), hipMemcpyHostToDevice, c->stream));
            JCI_CHECK(hipStreamSynchronize(c->stream));
            RCHECK(jk_gf_group(prec, bn, d_h, d_t, &hh, c->stream));
        }
        return 0;
    }
    if (!h->d_gft) {
        JCI_CHECK(hipMalloc(&h->d_gft, sizeof(GfTask) * JCI_MAXK));
        JCI_CHECK(hipMalloc(&h->d_gfh, sizeof(GfHeader) * JCI_MAXK));
        JCI_CHECK(hipHostMalloc(&h->h_gft, sizeof(GfTask) * JCI_MAXK, hipHostMallocDefault));
        JCI_CHECK(hipHostMalloc(&h->h_gfh, sizeof(GfHeader) * JCI_MAXK, hipHostMallocDefault));
        memset(h->gf_valid, 0, sizeof(h->gf_valid));
    }
    // vectors per pass: the intermediate E (La x k x N) stays within 8 GB
    const int Kc = (int)std::max(1.0, std::min((double)K, 8e9 / ((double)h->gf.La * N * esz)));
    for (int b = 0, ci = 0; b < K; b += Kc, ci++) {
        const int k = std::min(Kc, K - b);
        RCHECK(grow(&h->d_gfE, &h->gfE_bytes, (size_t)h->gf.La * jk_gf_lde((long long)k * N) * esz,
                    "the intermediate"));
        if (prec == 32)
            RCHECK(grow((void **)&h->d_gfX32, &h->gfX32_len, sizeof(float) * (size_t)k * N,
                        "FP32 staging"));
        GfTask t;
        GfHeader hh;
        memset(&t, 0, sizeof(t));
        memset(&hh, 0, sizeof(hh));
        t.g = h->gf;
        t.K = k;
        t.x64 = x + (size_t)b * ldx;
        t.ldx64 = ldx;
        t.x32 = prec == 32 ? h->d_gfX32 : nullptr;
        t.y = y + (size_t)b * ldy;
        t.ldy = ldy;
        t.E = h->d_gfE;
        t.lde = jk_gf_lde((long long)k * N);
        hh.n = 1;
        const int bn = prec == 64 ? h->gf_bn64 : h->gf_bn32;
        jk_gf_units(prec, bn, t, hh.total);
        if (!h->gf_valid[ci] || memcmp(&t, &h->h_gft[ci], sizeof(t)) ||
            memcmp(&hh, &h->h_gfh[ci], sizeof(hh))) {
            JCI_CHECK(hipStreamSynchronize(c->stream));   // earlier uploads of the slot done
            h->h_gft[ci] = t;
            h->h_gfh[ci] = hh;
            JCI_CHECK(hipMemcpyAsync(h->d_gft + ci, h->h_gft + ci, sizeof(t), hipMemcpyHostToDevice, c->stream));
            JCI_CHECK(hipMemcpyAsync(h->d_gfh + ci, h->h_gfh + ci, sizeof(hh), hipMemcpyHostToDevice, c->stream));
            h->gf_valid[ci] = 1;
        }
        RCHECK(jk_gf_group(prec, bn, h->d_gfh + ci, h->d_gft + ci, &hh, c->stream));
    }
    return 0;
}

static bool is_gf(int method) { return method == JCI_SIGMA_GEMM64 || method == JCI_SIGMA_GEMM32; }

// table sigma of nvec determinant-major vectors x; output determinant-major (ldy 0) or
// vector-major (vector k at y + k*ldy); s as for sigma_dev
static int fast_run(jci_ham *h, int method, int nvec, const double *x, double *y, long long ldy,
                    SigmaScratch *s)
{
    jci_ctx *c = h->sp->ctx;
    const long long N = h->sp->h.ndets;
    if (!h->have_fast) { jci_set_error("sigma: no excitation tables (JCI_DIRECT_TABLES)"); return -1; }
    const int prec = method == JCI_SIGMA_FAST64 ? 64 : 32;
    const void *xin = x;
    if (prec == 32) {
        float *xf;
        RCHECK(ensure_xf(h, N * nvec, s, &xf));
        RCHECK(jk_to_f32(x, xf, N * nvec, c->stream));
        xin = xf;
    }
    // one launch of the chosen width, or one per width class (fast_w == 3); two passes: per
    // class and orientation (set_orient)
    // alpha chunk per class unless one is set: FP64 1 for rows of up to 128 strings, 2 up
    // to 256, 4 above; FP32 2, 2, 4, 4 (LDS per work-group grows with the chunk and the
    // element size; measured at 16e/60o: FP64 7.44 s against 8.75 s with 2 and 8.87 s with
    // 8 throughout; FP32 5.86 s with 4 throughout against 6.02 s with 2/2/4/8, while at
    // 16e/20o 4 throughout took 42.2 ms against 38.3 ms)
    const int cls_a[4] = {prec == 64 ? 1 : 2, prec == 64 ? 1 : 2, prec == 64 ? 2 : 4, 4};
    const int A_set = h->fast[0].A;
    FastArgs a[8], t[8];
    int n = 1;
    if (h->fast_w == 3 && !h->fast_split) {
        n = 4;
        for (int i = 0; i < 4; i++) {
            a[i] = h->fcls[i];
            a[i].A = A_set > 0 ? A_set : cls_a[i];
        }
    } else if (h->fast_w == 3) {
        n = 8;
        for (int i = 0; i < 8; i++) {
            a[i] = h->fm[i];
            a[i].A = A_set > 0 ? A_set : cls_a[i % 4];
            t[i] = h->fc[i];
        }
    } else {
        a[0] = h->fast[h->fast_w];
        t[0] = h->fastT[h->fast_w];
    }
    for (int i = 0; i < n; i++) {
        a[i].ldy = t[i].ldy = ldy;
        a[i].xg = t[i].xg = h->fast_xg;
        a[i].cm = t[i].cm = h->fast_cm;
    }
    if (!h->fast_split) return jk_sigma_fast_multi(prec, nvec, a, n, xin, y, c->stream);
    const size_t need = (size_t)N * nvec * (prec == 64 ? sizeof(double) : sizeof(float));
    void *xT;
    if (s) {
        RCHECK(jci_scratch_carve(s, need, &xT, "the transposed input"));
    } else {
        RCHECK(grow(&h->d_xT, &h->xT_bytes, need, "the transposed input"));
        xT = h->d_xT;
    }
    return jk_sigma_fast_split(prec, nvec, h->sp->d_desc, a, n, t, n, h->d_tt_off, h->ntt, xin,
                               xT, y, c->stream);
}

// y = H x on device, x and y determinant-major [N][nvec]; s (may be NULL) supplies every
// device buffer the call needs beyond x and y (the allocation-free entry point)
static int sigma_dev(jci_ham *h, int method, int nvec, const double *x, double *y,
                     SigmaScratch *s = nullptr)
{
    jci_space *sp = h->sp;
    jci_ctx *c = sp->ctx;
    long long N = sp->h.ndets;
    float *xf = nullptr;
    if (nvec < 1 || nvec > JCI_MAXK) { jci_set_error("sigma: nvec must be 1..%d", JCI_MAXK); return -1; }
    method = resolve_method(h, method, 0);
    switch (method) {
    case JCI_SIGMA_PAIRS64:
        if (!h->d_val64 || h->upper) { jci_set_error("sigma: no full FP64 pair list"); return -1; }
        return jk_spmv(64, nvec, N, h->d_rowptr, h->d_col, h->d_val64, h->d_hd, x, y, c->stream);
    case JCI_SIGMA_PAIRS32:
        if (!h->d_val32 || h->upper) { jci_set_error("sigma: no full FP32 pair list"); return -1; }
        RCHECK(ensure_xf(h, N * nvec, s, &xf));
        RCHECK(jk_to_f32(x, xf, N * nvec, c->stream));
        return jk_spmv(32, nvec, N, h->d_rowptr, h->d_col, h->d_val32, h->d_hd, xf, y, c->stream);
    case JCI_SIGMA_PAIRS_SYM64:
        if (!h->d_val64 || !h->upper) { jci_set_error("sigma: no upper FP64 pair list"); return -1; }
        if (nvec == 1)
            return jk_spmv_sym(1, N, h->d_rowptr, h->d_col, h->d_val64, h->d_hd, x, y, c->stream);
        jci_set_error("sigma: symmetric pair list takes one vector per call");
        return -1;
    case JCI_SIGMA_DIRECT64:
        return jk_sigma_generic(64, nvec, sp->d_desc, N, sp->d_mask[0], sp->d_mask[1], c->d_binom,
                                h->d_m1, h->d_m2, h->d_hd, x, y, c->stream);
    case JCI_SIGMA_DIRECT32:
        RCHECK(ensure_xf(h, N * nvec, s, &xf));
        RCHECK(jk_to_f32(x, xf, N * nvec, c->stream));
        return jk_sigma_generic(32, nvec, sp->d_desc, N, sp->d_mask[0], sp->d_mask[1], c->d_binom,
                                h->d_m1, h->d_m2, h->d_hd, xf, y, c->stream);
    case JCI_SIGMA_FAST64:
    case JCI_SIGMA_FAST32:
        return fast_run(h, method, nvec, x, y, 0, s);
    case JCI_SIGMA_FCI64:
        if (!h->have_fci) { jci_set_error("sigma: single-block kernel not applicable"); return -1; }
        return jk_sigma_fci(64, nvec, h->fci, x, y, c->stream);
    case JCI_SIGMA_FCI32:
        if (!h->have_fci) { jci_set_error("sigma: single-block kernel not applicable"); return -1; }
        RCHECK(ensure_xf(h, N * nvec, s, &xf));
        RCHECK(jk_to_f32(x, xf, N * nvec, c->stream));
        return jk_sigma_fci(32, nvec, h->fci, xf, y, c->stream);
    case JCI_SIGMA_GEMM64:
    case JCI_SIGMA_GEMM32: {
        if (!h->have_gf) { jci_set_error("sigma: dense-string sigma not applicable"); return -1; }
        const int prec = method == JCI_SIGMA_GEMM64 ? 64 : 32;
        if (nvec == 1) return gf_run(h, prec, 1, x, N, y, N, s);
        double *t1;
        if (s) {
            RCHECK(jci_scratch_carve(s, sizeof(double) * 2 * (size_t)nvec * N, (void **)&t1,
                                     "vector staging"));
        } else {
            RCHECK(grow((void **)&h->d_gfT, &h->gfT_len, sizeof(double) * 2 * (size_t)nvec * N,
                        "vector staging"));
            t1 = h->d_gfT;
        }
        double *t2 = t1 + (size_t)nvec * N;
        RCHECK(jk_unpack(nvec, N, x, t1, N, c->stream));
        RCHECK(gf_run(h, prec, nvec, t1, N, t2, N, s));
        return jk_pack(nvec, N, t2, N, y, c->stream);
    }
    default:
        jci_set_error("sigma: unknown method %d", method);
        return -1;
    }
}

JCI_API int jci_sigma_dev(jci_ham *h, int method, int nvec, const double *x_dev, double *y_dev)
{
    JCI_CHECK(hipSetDevice(h->sp->ctx->device));
    return sigma_dev(h, method, nvec, x_dev, y_dev);
}

int jci_ham_device(const jci_ham *h) { return h->sp->ctx->device; }

int jci_sigma_dev_scratch(jci_ham *h, int method, int nvec, const double *x_dev, double *y_dev,
                          SigmaScratch *s)
{
    JCI_CHECK(hipSetDevice(h->sp->ctx->device));
    return sigma_dev(h, method, nvec, x_dev, y_dev, s);
}

JCI_API int jci_sigma(jci_ham *h, int method, int nvec, const double *x, double *y)
{
    jci_ctx *c = h->sp->ctx;
    JCI_CHECK(hipSetDevice(c->device));
    long long N = h->sp->h.ndets;
    double *dx, *dy;
    JCI_CHECK(hipMalloc(&dx, sizeof(double) * N * nvec));
    JCI_CHECK(hipMalloc(&dy, sizeof(double) * N * nvec));
    JCI_CHECK(hipMemcpyAsync(dx, x, sizeof(double) * N * nvec, hipMemcpyHostToDevice, c->stream));
    int rc = sigma_dev(h, method, nvec, dx, dy);
    if (rc == 0) {
        JCI_CHECK(hipMemcpyAsync(y, dy, sizeof(double) * N * nvec, hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
    }
    hipFree(dx);
    hipFree(dy);
    return rc;
}

// ---- small dense eigensolver (cyclic Jacobi, FP64) ------------------------------------------
// eigenvalues ascending; eigenvector i in vec[i*n + 0..n-1]
void jci_sym_eig(int n, const double *a, double *eval, double *vec);   // symeig.cpp

// subspace eigensolver (eigenvalues ascending, vectors row-major with a fixed sign)
static void jacobi_eig(int n, const double *a_in, double *eval, double *vec)
{
    jci_sym_eig(n, a_in, eval, vec);
}

JCI_API int jci_gram(jci_ctx *c, int m, int w, long long n, const double *V, const double *W,
                     double *G)
{
    if (m < 1 || w < 1 || m > JCI_MAXBASIS || w > JCI_MAXBASIS || n < 1) {
        jci_set_error("gram: dimensions out of range");
        return -1;
    }
    JCI_CHECK(hipSetDevice(c->device));
    double *dV, *dW;
    JCI_CHECK(hipMalloc(&dV, sizeof(double) * m * n));
    JCI_CHECK(hipMalloc(&dW, sizeof(double) * w * n));
    JCI_CHECK(hipMemcpyAsync(dV, V, sizeof(double) * m * n, hipMemcpyHostToDevice, c->stream));
    JCI_CHECK(hipMemcpyAsync(dW, W, sizeof(double) * w * n, hipMemcpyHostToDevice, c->stream));
    int rc = jk_gram(m, w, n, dV, n, dW, n, c->d_partial, JCI_PARTIAL_BLOCKS, c->d_small, c->stream);
    if (rc == 0) {
        JCI_CHECK(hipMemcpyAsync(G, c->d_small, sizeof(double) * m * w, hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
    }
    hipFree(dV);
    hipFree(dW);
    return rc;
}

JCI_API int jci_bench_sigma(jci_ham *h, int method, int nvec, int reps, double *ms)
{
    jci_ctx *c = h->sp->ctx;
    JCI_CHECK(hipSetDevice(c->device));
    long long N = h->sp->h.ndets;
    if (reps < 1 || nvec < 1 || nvec > JCI_MAXK) { jci_set_error("bench_sigma: bad arguments"); return -1; }
    double *x, *y;
    JCI_CHECK(hipMalloc(&x, sizeof(double) * N * nvec));
    JCI_CHECK(hipMalloc(&y, sizeof(double) * N * nvec));
    int rc = jk_fill_random(N * nvec, 12345, x, c->stream);
    if (!rc) rc = sigma_dev(h, method, nvec, x, y);       // warm-up launch
    hipEvent_t e0, e1;
    JCI_CHECK(hipEventCreate(&e0));
    JCI_CHECK(hipEventCreate(&e1));
    if (!rc) {
        JCI_CHECK(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps && !rc; r++) rc = sigma_dev(h, method, nvec, x, y);
        JCI_CHECK(hipEventRecord(e1, c->stream));
        JCI_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        JCI_CHECK(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(x);
    hipFree(y);
    return rc;
}

// Times every available sigma variant (pair list, matrix-free with tile widths 64/128/256,
// one-thread-per-row) for nvec vectors and makes the fastest FP64 and FP32 ones the
// defaults of JCI_SIGMA_AUTO.  times[] (may be NULL) receives ms per launch indexed by
// JCI_SIGMA_* for the chosen tile width (0 where not available).
JCI_API int jci_ham_autotune(jci_ham *h, int nvec, int reps, int *best64, int *best32,
                             double *times)
{
    double t[JCI_SIGMA_GEMM32 + 1] = {0};
    double best_w_ms = 1e300;
    int best_w = h->fast_w;
    if (h->have_fast) {
        // beta tile width x alpha chunk of the table kernel
        int best_a = 0, best_split = h->fast_split;
        const int chunks[4] = {0, 1, 2, 4};
        for (int sp = 0; sp < 3; sp++)
            for (int wi = 0; wi < 4; wi++)
                for (int ai = 0; ai < 4; ai++) {
                    if (h->fast_split != sp) {
                        h->fast_split = sp;
                        RCHECK(set_orient(h));
                    }
                    h->fast_w = wi;
                    for (int w2 = 0; w2 < 3; w2++) h->fast[w2].A = chunks[ai];
                    double ms = 0;
                    RCHECK(jci_bench_sigma(h, JCI_SIGMA_FAST64, nvec, reps, &ms));
                    if (ms < best_w_ms) {
                        best_w_ms = ms; best_w = wi; best_a = chunks[ai]; best_split = sp;
                    }
                }
        h->fast_w = best_w;
        if (h->fast_split != best_split) {
            h->fast_split = best_split;
            RCHECK(set_orient(h));
        }
        for (int w2 = 0; w2 < 3; w2++) h->fast[w2].A = best_a;
        t[JCI_SIGMA_FAST64] = best_w_ms;
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_FAST32, nvec, reps, &t[JCI_SIGMA_FAST32]));
    }
    if (h->have_fci) {
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_FCI64, nvec, reps, &t[JCI_SIGMA_FCI64]));
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_FCI32, nvec, reps, &t[JCI_SIGMA_FCI32]));
    }
    if (h->have_gf) {
        // E tile widths: keep the fastest per precision
        const int w64[2] = {16, 32}, w32[2] = {32, 64};
        double b64 = 1e300, b32 = 1e300;
        int k64 = h->gf_bn64, k32 = h->gf_bn32;
        for (int i = 0; i < 2; i++) {
            double ms = 0;
            h->gf_bn64 = w64[i];
            RCHECK(jci_bench_sigma(h, JCI_SIGMA_GEMM64, nvec, reps, &ms));
            if (ms < b64) { b64 = ms; k64 = w64[i]; }
            h->gf_bn32 = w32[i];
            RCHECK(jci_bench_sigma(h, JCI_SIGMA_GEMM32, nvec, reps, &ms));
            if (ms < b32) { b32 = ms; k32 = w32[i]; }
        }
        h->gf_bn64 = k64;
        h->gf_bn32 = k32;
        t[JCI_SIGMA_GEMM64] = b64;
        t[JCI_SIGMA_GEMM32] = b32;
    }
    if (h->d_val64 && !h->upper) RCHECK(jci_bench_sigma(h, JCI_SIGMA_PAIRS64, nvec, reps, &t[JCI_SIGMA_PAIRS64]));
    if (h->d_val32 && !h->upper) RCHECK(jci_bench_sigma(h, JCI_SIGMA_PAIRS32, nvec, reps, &t[JCI_SIGMA_PAIRS32]));
    if (!h->have_fast && !h->d_val64 && !h->d_val32) {
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_DIRECT64, nvec, 1, &t[JCI_SIGMA_DIRECT64]));
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_DIRECT32, nvec, 1, &t[JCI_SIGMA_DIRECT32]));
    }
    int b64 = 0, b32 = 0;
    const int c64[] = {JCI_SIGMA_PAIRS64, JCI_SIGMA_FAST64, JCI_SIGMA_DIRECT64, JCI_SIGMA_FCI64,
                       JCI_SIGMA_GEMM64};
    const int c32[] = {JCI_SIGMA_PAIRS32, JCI_SIGMA_FAST32, JCI_SIGMA_DIRECT32, JCI_SIGMA_FCI32,
                       JCI_SIGMA_GEMM32};
    for (int m : c64) if (t[m] > 0 && (b64 == 0 || t[m] < t[b64])) b64 = m;
    for (int m : c32) if (t[m] > 0 && (b32 == 0 || t[m] < t[b32])) b32 = m;
    h->tuned64 = b64;
    h->tuned32 = b32;
    if (best64) *best64 = b64;
    if (best32) *best32 = b32;
    if (times) for (int m = 0; m <= JCI_SIGMA_GEMM32; m++) times[m] = t[m];
    return 0;
}

JCI_API int jci_ham_fast_tiles(jci_ham *h, int *width, int *chunk)
{
    if (!h->have_fast) { jci_set_error("ham: no excitation tables"); return -1; }
    if (width && *width >= 0) {
        int wi = *width == 64 ? 0 : *width == 128 ? 1 : *width == 256 ? 2 : *width == 0 ? 3 : -1;
        if (wi < 0) {
            jci_set_error("ham: tile width %d not available (64, 128, 256, or 0: per block)", *width);
            return -1;
        }
        h->fast_w = wi;
    }
    if (chunk && *chunk >= 0) {
        if (*chunk != 0 && *chunk != 1 && *chunk != 2 && *chunk != 4 && *chunk != 8) {
            jci_set_error("ham: alpha chunk %d not available (0, 1, 2, 4, 8)", *chunk);
            return -1;
        }
        for (int w2 = 0; w2 < 3; w2++) h->fast[w2].A = *chunk;
    }
    if (width) *width = h->fast_w == 3 ? 0 : h->fast[h->fast_w].W;
    if (chunk) *chunk = h->fast[h->fast_w == 3 ? 0 : h->fast_w].A;
    return 0;
}

JCI_API int jci_ham_fast_split(jci_ham *h, int *split)
{
    if (!h->have_fast) { jci_set_error("ham: no excitation tables"); return -1; }
    if (split && *split > 3) { jci_set_error("ham: split mode %d not 0 .. 3", *split); return -1; }
    if (split && *split >= 0 && *split != h->fast_split) {
        h->fast_split = *split;
        RCHECK(set_orient(h));
    }
    if (split) *split = h->fast_split;
    return 0;
}

JCI_API int jci_ham_fast_xcd(jci_ham *h, int *group, int *chunk_major)
{
    if (!h->have_fast) { jci_set_error("ham: no excitation tables"); return -1; }
    if (group && *group >= 0) {
        if (*group > 4096) { jci_set_error("ham: XCD run of %d tiles too long", *group); return -1; }
        h->fast_xg = *group;
    }
    if (chunk_major && *chunk_major >= 0) h->fast_cm = *chunk_major ? 1 : 0;
    if (group) *group = h->fast_xg;
    if (chunk_major) *chunk_major = h->fast_cm;
    return 0;
}

JCI_API int jci_ham_gemm_tiles(jci_ham *h, int *bn64, int *bn32)
{
    if (!h->have_gf) { jci_set_error("ham: no dense-string sigma"); return -1; }
    if (bn64 && *bn64 > 0) {
        if (!jk_gf_bn_valid(64, *bn64)) { jci_set_error("ham: FP64 tile width %d not available", *bn64); return -1; }
        h->gf_bn64 = *bn64;
    }
    if (bn32 && *bn32 > 0) {
        if (!jk_gf_bn_valid(32, *bn32)) { jci_set_error("ham: FP32 tile width %d not available", *bn32); return -1; }
        h->gf_bn32 = *bn32;
    }
    if (bn64) *bn64 = h->gf_bn64;
    if (bn32) *bn32 = h->gf_bn32;
    return 0;
}

JCI_API int jci_ham_fast_width(const jci_ham *h)
{
    return h->have_fast ? (h->fast_w == 3 ? 0 : h->fast[h->fast_w].W) : 0;
}

JCI_API int jci_bench_copy(jci_ctx *c, long long bytes, int reps, double *ms)
{
    JCI_CHECK(hipSetDevice(c->device));
    if (reps < 1 || bytes < 32) { jci_set_error("bench_copy: bad arguments"); return -1; }
    bytes = bytes / 32 * 32;
    void *a, *b;
    JCI_CHECK(hipMalloc(&a, bytes));
    JCI_CHECK(hipMalloc(&b, bytes));
    int rc = jk_fill_random(bytes / 8, 7, (double *)a, c->stream);
    if (!rc) rc = jk_copy(bytes, a, b, c->stream);
    hipEvent_t e0, e1;
    JCI_CHECK(hipEventCreate(&e0));
    JCI_CHECK(hipEventCreate(&e1));
    if (!rc) {
        JCI_CHECK(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps && !rc; r++) rc = jk_copy(bytes, a, b, c->stream);
        JCI_CHECK(hipEventRecord(e1, c->stream));
        JCI_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        JCI_CHECK(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return rc;
}

void jci_jacobi_eig(int n, const double *a, double *eval, double *vec)
{
    jacobi_eig(n, a, eval, vec);
}

int jci_gdk_restart(int m, const double *G, const double *Yf, int kc, int mp, int np,
                    const double *Yp, int kmax, double *C, double *Gc)
{
    std::vector<double> cols;   // column-major, m rows
    int n = 0;
    kc = std::min(kc, kmax);
    for (int j = 0; j < kc; j++, n++) cols.insert(cols.end(), Yf + (size_t)j * m, Yf + (size_t)j * m + m);
    std::vector<double> v(m);
    for (int j = 0; j < np && n < kmax; j++) {
        for (int i = 0; i < m; i++) v[i] = i < mp ? Yp[(size_t)j * mp + i] : 0.0;
        for (int pass = 0; pass < 2; pass++)
            for (int q = 0; q < n; q++) {
                const double *c = &cols[(size_t)q * m];
                double d = 0.0;
                for (int i = 0; i < m; i++) d += c[i] * v[i];
                for (int i = 0; i < m; i++) v[i] -= d * c[i];
            }
        double nr = 0.0;
        for (int i = 0; i < m; i++) nr += v[i] * v[i];
        nr = sqrt(nr);
        if (nr < 1e-8) continue;
        for (int i = 0; i < m; i++) v[i] /= nr;
        cols.insert(cols.end(), v.begin(), v.end());
        n++;
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) C[(size_t)i * n + j] = cols[(size_t)j * m + i];
    std::vector<double> GC((size_t)m * n);   // G C, row-major
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double t = 0.0;
            for (int l = 0; l < m; l++) t += G[(size_t)i * m + l] * cols[(size_t)j * m + l];
            GC[(size_t)i * n + j] = t;
        }
    for (int a = 0; a < n; a++)
        for (int b = 0; b <= a; b++) {
            double ab = 0.0, ba = 0.0;
            for (int i = 0; i < m; i++) {
                ab += cols[(size_t)a * m + i] * GC[(size_t)i * n + b];
                ba += cols[(size_t)b * m + i] * GC[(size_t)i * n + a];
            }
            Gc[(size_t)a * n + b] = Gc[(size_t)b * n + a] = 0.5 * (ab + ba);
        }
    return n;
}

JCI_API int jci_jacobi(int n, const double *a, double *eval, double *vec)
{
    if (n < 1) return -1;
    jacobi_eig(n, a, eval, vec);
    return 0;
}

// ---- Davidson ------------------------------------------------------------------------------
JCI_API void jci_solve_defaults(jci_solve_params *p)
{
    memset(p, 0, sizeof(*p));
    p->maxiter = 200;
    p->krymin = 4;
    p->krymax = 16;
    p->nroots = 1;
    p->refdim = 64;
    p->restol = 1e-6;
    p->precision = JCI_PREC_FP64;
    p->mixed_switch = 1e-4;
    p->guess = JCI_GUESS_BLOCK;
}

struct Work {
    SolveOp *op;
    jci_ctx *c;
    long long N, ld;
    int kb;                      // vectors per block: new directions per iteration, sigma chunk
    double *V, *S, *T, *X, *Y;   // V,S: krymax vectors; T: kb vectors; X,Y: [N][kb]
    double sigma_time;
    long long ns64, ns32;
};

// S[first .. first+cnt) = H V[first .. first+cnt), in chunks of w.kb
static int sigma_block(Work &w, int first, int cnt, int method)
{
    double t0 = now_s();
    for (int b = first; b < first + cnt; b += w.kb) {
        int k = std::min(w.kb, first + cnt - b);
        if (w.op->sigma_cols) {
            int rc = w.op->sigma_cols(w.op->self, method, k, w.V + (size_t)b * w.ld, w.ld,
                                      w.S + (size_t)b * w.ld, w.ld, w.X);
            if (rc < 0) return rc;
            if (rc == 0) continue;
        }
        if (!w.Y && hipMalloc(&w.Y, sizeof(double) * (size_t)w.ld * w.kb) != hipSuccess) {
            jci_set_error("solve: out of device memory for the sigma staging block");
            return -3;
        }
        RCHECK(jk_pack(k, w.N, w.V + (size_t)b * w.ld, w.ld, w.X, w.c->stream));
        RCHECK(w.op->sigma(w.op->self, method, k, w.X, w.Y));
        RCHECK(jk_unpack(k, w.N, w.Y, w.S + (size_t)b * w.ld, w.ld, w.c->stream));
    }
    JCI_CHECK(hipStreamSynchronize(w.c->stream));
    w.sigma_time += now_s() - t0;
    return 0;
}

static int gram_host(Work &w, int m, const double *A, int n, const double *B, double *out)
{
    RCHECK(jk_gram(m, n, w.N, A, w.ld, B, w.ld, w.c->d_partial, JCI_PARTIAL_BLOCKS, w.c->d_small,
                   w.c->stream));
    JCI_CHECK(hipMemcpyAsync(w.c->h_small, w.c->d_small, sizeof(double) * m * n,
                             hipMemcpyDeviceToHost, w.c->stream));
    JCI_CHECK(hipStreamSynchronize(w.c->stream));
    memcpy(out, w.c->h_small, sizeof(double) * m * n);
    return 0;
}

// upload a small matrix to device scratch (after the device-side uses of d_small finished)
static int upload_small(Work &w, const double *src, int n, double **dptr)
{
    JCI_CHECK(hipStreamSynchronize(w.c->stream));
    memcpy(w.c->h_small, src, sizeof(double) * n);
    JCI_CHECK(hipMemcpyAsync(w.c->d_small, w.c->h_small, sizeof(double) * n, hipMemcpyHostToDevice,
                             w.c->stream));
    *dptr = w.c->d_small;
    return 0;
}

// orthonormalize T[0..k) against V[0..m) (two classical Gram-Schmidt passes) and among
// themselves (symmetric orthonormalization, dropping dependent directions); returns the
// number of vectors kept in *kept (moved to the front of T)
static int orthonormalize_new(Work &w, int m, int k, int *kept)
{
    std::vector<double> cmat((size_t)m * k), nb((size_t)m * k);
    double *dY;
    for (int pass = 0; pass < 2 && m > 0; pass++) {
        RCHECK(gram_host(w, m, w.V, k, w.T, cmat.data()));
        for (size_t i = 0; i < cmat.size(); i++) nb[i] = -cmat[i];
        RCHECK(upload_small(w, nb.data(), m * k, &dY));
        RCHECK(jk_combine(m, k, w.N, w.V, w.ld, dY, k, w.T, w.ld, 1.0, w.c->stream));
    }
    std::vector<double> g((size_t)k * k), ev(k), U((size_t)k * k);
    RCHECK(gram_host(w, k, w.T, k, w.T, g.data()));
    for (int i = 0; i < k; i++)
        for (int j = 0; j < i; j++) g[i * k + j] = g[j * k + i] = 0.5 * (g[i * k + j] + g[j * k + i]);
    jacobi_eig(k, g.data(), ev.data(), U.data());
    double emax = 0.0;
    for (int i = 0; i < k; i++) emax = std::max(emax, ev[i]);
    // keep directions with eigenvalue above a relative threshold, largest first
    std::vector<double> Y;
    int nk = 0;
    for (int i = k - 1; i >= 0; i--) {
        if (ev[i] <= 1e-12 * emax || ev[i] <= 1e-24) continue;
        nk++;
    }
    Y.assign((size_t)k * nk, 0.0);
    int col = 0;
    for (int i = k - 1; i >= 0 && col < nk; i--) {
        if (ev[i] <= 1e-12 * emax || ev[i] <= 1e-24) continue;
        double s = 1.0 / sqrt(ev[i]);
        for (int r = 0; r < k; r++) Y[(size_t)r * nk + col] = U[(size_t)i * k + r] * s;
        col++;
    }
    if (nk > 0) {
        RCHECK(upload_small(w, Y.data(), k * nk, &dY));
        RCHECK(jk_rotate_inplace(k, nk, w.N, w.T, w.ld, dY, w.c->stream));
    }
    *kept = nk;
    return 0;
}

// indices of the nb lowest diagonal elements, ascending by (value, index): each chunk's nb
// lowest selected on the device (jk_lowest_chunks), the candidates merged here; past
// JCI_GUESS_CHUNK the whole diagonal is sorted here
static int lowest_diag(jci_ctx *c, const double *d_hd, long long N, int nb, long long *idx)
{
    if (nb < 1 || nb > N) { jci_set_error("lowest_diag: %d of %lld", nb, N); return -1; }
    std::vector<double> hv;
    std::vector<long long> hi;
    if (nb <= JCI_GUESS_CHUNK) {
        const size_t ncand = (size_t)((N + JCI_GUESS_CHUNK - 1) / JCI_GUESS_CHUNK) * nb;
        void *cand = nullptr;
        JCI_CHECK(hipMallocAsync(&cand, (sizeof(double) + sizeof(long long)) * ncand, c->stream));
        double *d_cv = (double *)cand;
        long long *d_ci = (long long *)(d_cv + ncand);
        hv.resize(ncand);
        hi.resize(ncand);
        int rc = jk_lowest_chunks(N, d_hd, nb, d_cv, d_ci, c->stream);
        if (!rc && (hipMemcpyAsync(hv.data(), d_cv, sizeof(double) * ncand, hipMemcpyDeviceToHost,
                                   c->stream) != hipSuccess ||
                    hipMemcpyAsync(hi.data(), d_ci, sizeof(long long) * ncand, hipMemcpyDeviceToHost,
                                   c->stream) != hipSuccess ||
                    hipStreamSynchronize(c->stream) != hipSuccess))
            rc = -2;
        hipFreeAsync(cand, c->stream);
        if (rc) return rc;
    } else {
        hv.resize(N);
        hi.resize(N);
        JCI_CHECK(hipMemcpyAsync(hv.data(), d_hd, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
        for (long long i = 0; i < N; i++) hi[i] = i;
    }
    std::vector<size_t> ord(hv.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = i;
    std::partial_sort(ord.begin(), ord.begin() + nb, ord.end(), [&](size_t a, size_t b) {
        return hv[a] < hv[b] || (hv[a] == hv[b] && hi[a] < hi[b]);
    });
    for (int i = 0; i < nb; i++) idx[i] = hi[ord[i]];
    return 0;
}

JCI_API int jci_ham_lowest_diag(jci_ham *h, int nb, long long *idx)
{
    JCI_CHECK(hipSetDevice(h->sp->ctx->device));
    return lowest_diag(h->sp->ctx, h->d_hd, h->sp->h.ndets, nb, idx);
}

// the initial block: its nb determinants (the lowest diagonal elements, or the leading
// ones) and the lowest nguess eigenvectors of H on them (vector j at U + j*nb)
static int guess_compute(SolveOp *op, jci_ctx *c, const jci_solve_params *p, int nguess,
                         std::vector<long long> &idx, std::vector<double> &U)
{
    const long long N = op->N;
    int nb = (int)std::min<long long>(p->refdim, N);
    if (nb < nguess) nb = (int)std::min<long long>(nguess, N);
    idx.assign(nb, 0);
    if (p->guess == JCI_GUESS_DIAG) {
        RCHECK(lowest_diag(c, op->d_hd, N, nb, idx.data()));
    } else {
        for (int i = 0; i < nb; i++) idx[i] = i;
    }
    // device scratch in one block: indices, masks, the block of H
    void *scratch = nullptr;
    JCI_CHECK(hipMallocAsync(&scratch, sizeof(long long) * nb + sizeof(uint64_t) * 2 * nb +
                                           sizeof(double) * (size_t)nb * nb, c->stream));
    long long *d_idx = (long long *)scratch;
    uint64_t *d_m = (uint64_t *)(d_idx + nb);
    double *d_H = (double *)(d_m + 2 * nb);
    std::vector<double> H((size_t)nb * nb), ev(nb), Uf((size_t)nb * nb);
    int rc = 0;
    if (hipMemcpyAsync(d_idx, idx.data(), sizeof(long long) * nb, hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        rc = -2;
    if (!rc) rc = op->masks(op->self, nb, d_idx, d_m, d_m + nb);
    if (!rc) rc = jk_dense_block(nb, d_m, d_m + nb, 0, op->d_m1, op->d_m2, d_H, c->stream);
    if (!rc && (hipMemcpyAsync(H.data(), d_H, sizeof(double) * nb * nb, hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = -2;
    hipFreeAsync(scratch, c->stream);
    if (rc) return rc;
    jacobi_eig(nb, H.data(), ev.data(), Uf.data());
    U.assign(Uf.begin(), Uf.begin() + (size_t)nb * nguess);
    return 0;
}

// V[:, 0..nguess) = 0, then the guess vectors scattered to their determinants, on stream st;
// d_scr holds (8 + 8 nguess) nb bytes
static int guess_apply(const std::vector<long long> &idx, const std::vector<double> &U,
                       int nguess, double *V, long long ld, void *d_scr, hipStream_t st)
{
    const int nb = (int)idx.size();
    long long *d_idx = (long long *)d_scr;
    double *d_U = (double *)(d_idx + nb);
    JCI_CHECK(hipMemcpyAsync(d_idx, idx.data(), sizeof(long long) * nb, hipMemcpyHostToDevice, st));
    JCI_CHECK(hipMemcpyAsync(d_U, U.data(), sizeof(double) * nb * nguess, hipMemcpyHostToDevice, st));
    JCI_CHECK(hipMemsetAsync(V, 0, sizeof(double) * ld * nguess, st));
    RCHECK(jk_scatter_guess(nb, nguess, d_idx, d_U, V, ld, st));
    return 0;
}

static int guess_block(Work &w, const jci_solve_params *p, int nguess)
{
    std::vector<long long> idx;
    std::vector<double> U;
    RCHECK(guess_compute(w.op, w.c, p, nguess, idx, U));
    void *scr = nullptr;
    JCI_CHECK(hipMalloc(&scr, (sizeof(long long) + sizeof(double) * nguess) * idx.size()));
    int rc = guess_apply(idx, U, nguess, w.V, w.ld, scr, w.c->stream);
    if (!rc && hipStreamSynchronize(w.c->stream) != hipSuccess) rc = -2;
    hipFree(scr);
    return rc;
}

void jci_subspace_sizes(const jci_solve_params *p, long long N, int *ntrack, int *krymin,
                        int *krymax)
{
    const int kmax = (int)std::min<long long>(p->krymax, N);
    const int nt = (int)std::min(std::min<long long>(p->nroots + p->nbuffer, JCI_MAXROOTS),
                                 std::min<long long>(N, std::max<long long>(p->nroots, kmax / 2)));
    *krymax = kmax;
    *ntrack = nt;
    *krymin = std::min(std::max(p->krymin, nt), kmax);
}

int jci_solve_core(SolveOp *op, const jci_solve_params *p, const double *guess, double *evals,
                   double *evecs, jci_solve_stats *stats, int nout)
{
    double t_start = now_s();
    jci_ctx *c = op->ctx;
    JCI_CHECK(hipSetDevice(c->device));
    long long N = op->N;
    int nroots = p->nroots, krymax = std::min<long long>(p->krymax, N);
    if (nroots < 1 || nroots > JCI_MAXROOTS || nroots > N) { jci_set_error("solve: nroots out of range"); return -1; }
    if (krymax > JCI_MAXBASIS) { jci_set_error("solve: krymax above %d", JCI_MAXBASIS); return -1; }
    if (p->nbuffer < 0) { jci_set_error("solve: nbuffer must be >= 0"); return -1; }
    // tracked Ritz pairs: the requested roots plus the buffer ones (space permitting)
    const long long nt_cap = std::max<long long>(nroots, krymax / 2);
    const int ntrack = (int)std::min(std::min<long long>(nroots + p->nbuffer, JCI_MAXROOTS),
                                     std::min<long long>(N, nt_cap));
    int krymin = std::max(p->krymin, ntrack);
    if (krymin > krymax) krymin = krymax;
    if (krymin < nroots) { jci_set_error("solve: krymax below nroots"); return -1; }
    Work w{};
    w.op = op; w.c = c; w.N = N;
    w.ld = (N + 63) / 64 * 64;
    // block of the new directions of one iteration (ntrack at most) and of the sigma chunks:
    // the tracked pairs, at least four (JCI_MAXK at most); above 1e8 determinants exactly
    // the tracked pairs (device memory)
    w.kb = std::min(JCI_MAXK, std::max(ntrack, N > 100000000LL ? 1 : 4));
    size_t vb = sizeof(double) * (size_t)w.ld;
    // the determinant-major output staging w.Y is allocated at its first use (methods that
    // write vector-major columns need none)
    if (hipMalloc(&w.V, vb * krymax) != hipSuccess || hipMalloc(&w.S, vb * krymax) != hipSuccess ||
        hipMalloc(&w.T, vb * w.kb) != hipSuccess || hipMalloc(&w.X, vb * w.kb) != hipSuccess) {
        hipFree(w.V); hipFree(w.S); hipFree(w.T); hipFree(w.X); hipFree(w.Y);
        jci_set_error("solve: out of device memory for %d basis vectors of %lld", krymax, N);
        return -3;
    }
    int rc = 0;
    int mixed = p->precision == JCI_PREC_MIXED;
    int phase32 = mixed;
    double best32 = 1e300;      // FP32 phase: smallest largest-residual so far and when
    int it_best32 = 0;
    int m64 = op->resolve(op->self, p->sigma_fp64, 0), m32 = op->resolve(op->self, p->sigma_fp32, 1);
    auto cur_method = [&]() { return phase32 ? m32 : m64; };
    auto count_sigma = [&](int n) { if (phase32) w.ns32 += n; else w.ns64 += n; };
    int m = 0, it = 0, converged = 0;
    std::vector<double> G, Gk, Gblk, theta(JCI_MAXBASIS), Yf, rn(JCI_MAXROOTS, 0.0);
    int gvalid = 0;
    // Ritz vectors of the tracked pairs at the previous iteration (GD+k restart), np vectors
    // of mp rows in the current basis (zero below mp)
    std::vector<double> Yp;
    int mp = 0, np = 0;
    double *dY;
    // initial space
    if (p->guess == JCI_GUESS_USER) {
        if (!guess) { jci_set_error("solve: user guess requested without vectors"); rc = -1; goto done; }
        // the caller supplies p->krymin vectors; the effective krymin can be larger (tracked
        // buffer pairs), and the vectors beyond the caller's come from the block guess
        const int nsup = std::max(0, std::min(p->krymin, krymin));
        for (int i = 0; i < nsup; i++) {
            if (hipMemcpyAsync(w.T + (size_t)(i % w.kb) * w.ld, guess + (size_t)i * N,
                               sizeof(double) * N, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
                rc = -2; goto done;
            }
            if (i % w.kb == w.kb - 1 || i == nsup - 1) {
                int k = i % w.kb + 1, kept;
                if ((rc = orthonormalize_new(w, m, k, &kept))) goto done;
                if (hipMemcpyAsync(w.V + (size_t)m * w.ld, w.T, vb * kept, hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess) { rc = -2; goto done; }
                m += kept;
            }
        }
        if (m < krymin) {
            // block-guess vectors (staged in S, which holds no sigma yet) fill the rest,
            // orthonormalised against the caller's
            std::vector<long long> idx;
            std::vector<double> U;
            if ((rc = guess_compute(op, c, p, krymin, idx, U))) goto done;
            void *scr = nullptr;
            if (hipMalloc(&scr, (sizeof(long long) + sizeof(double) * krymin) * idx.size()) != hipSuccess) {
                jci_set_error("solve: out of device memory for the guess block");
                rc = -3; goto done;
            }
            rc = guess_apply(idx, U, krymin, w.S, w.ld, scr, c->stream);
            if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = -2;
            hipFree(scr);
            if (rc) goto done;
            for (int i = 0; i < krymin && m < krymin;) {
                const int k = std::min(w.kb, krymin - i);
                int kept;
                if (hipMemcpyAsync(w.T, w.S + (size_t)i * w.ld, vb * k, hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess) { rc = -2; goto done; }
                if ((rc = orthonormalize_new(w, m, k, &kept))) goto done;
                kept = std::min(kept, krymin - m);
                if (hipMemcpyAsync(w.V + (size_t)m * w.ld, w.T, vb * kept, hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess) { rc = -2; goto done; }
                m += kept;
                i += k;
            }
        }
    } else if (p->guess == JCI_GUESS_UNIT) {
        // unit vectors on determinants 0..krymin-1 (mpi_source/pdavidson.c:418-431)
        if (hipMemsetAsync(w.V, 0, vb * krymin, c->stream) != hipSuccess) { rc = -2; goto done; }
        static const double one = 1.0;
        for (int i = 0; i < krymin; i++)
            if (hipMemcpyAsync(w.V + (size_t)i * w.ld + i, &one, sizeof(double), hipMemcpyHostToDevice,
                               c->stream) != hipSuccess) { rc = -2; goto done; }
        m = krymin;
    } else {
        if ((rc = guess_block(w, p, krymin))) goto done;
        m = krymin;
    }
    if (m < nroots) { jci_set_error("solve: initial space has fewer vectors than roots"); rc = -1; goto done; }
    if ((rc = sigma_block(w, 0, m, cur_method()))) goto done;
    count_sigma(m);
    // the projected matrix V^T S is kept across iterations (leading dimension JCI_MAXBASIS):
    // columns [0, gvalid) are current, so an iteration computes only V^T S_new for the
    // vectors appended since; a collapse leaves diag(theta), a full sigma recompute 0
    Gk.assign((size_t)JCI_MAXBASIS * JCI_MAXBASIS, 0.0);
    for (it = 1; it <= p->maxiter; it++) {
        if (gvalid > m) gvalid = 0;
        if (gvalid < m) {
            const int nn = m - gvalid;
            Gblk.assign((size_t)m * nn, 0.0);
            if ((rc = gram_host(w, m, w.V, nn, w.S + (size_t)gvalid * w.ld, Gblk.data()))) goto done;
            for (int i = 0; i < m; i++)
                for (int j = 0; j < nn; j++) {
                    const int J = gvalid + j;
                    const double v = Gblk[(size_t)i * nn + j];
                    if (i < gvalid) {
                        Gk[(size_t)i * JCI_MAXBASIS + J] = Gk[(size_t)J * JCI_MAXBASIS + i] = v;
                    } else if (i <= J) {
                        // both new: symmetrise V_i.S_J and V_J.S_i
                        const double u = 0.5 * (v + Gblk[(size_t)J * nn + (i - gvalid)]);
                        Gk[(size_t)i * JCI_MAXBASIS + J] = Gk[(size_t)J * JCI_MAXBASIS + i] = u;
                    }
                }
            gvalid = m;
        }
        G.assign((size_t)m * m, 0.0);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) G[(size_t)i * m + j] = Gk[(size_t)i * JCI_MAXBASIS + j];
        Yf.assign((size_t)m * m, 0.0);
        jacobi_eig(m, G.data(), theta.data(), Yf.data());
        // residuals of the lowest ntrack Ritz pairs -> T (preconditioned), norms
        const int nt = std::min(ntrack, m);
        std::vector<double> Ysel((size_t)m * nt);
        for (int i = 0; i < m; i++)
            for (int r = 0; r < nt; r++) Ysel[(size_t)i * nt + r] = Yf[(size_t)r * m + i];
        double *d_th;
        // Ysel and theta share the small scratch: Ysel at [0, m*nt), theta after it
        std::vector<double> packed(Ysel);
        packed.insert(packed.end(), theta.begin(), theta.begin() + nt);
        if ((rc = upload_small(w, packed.data(), (int)packed.size(), &dY))) goto done;
        d_th = dY + (size_t)m * nt;
        if ((rc = jk_residual(m, nt, N, w.V, w.S, w.ld, dY, d_th, op->d_hd, w.T,
                              c->d_partial, JCI_PARTIAL_BLOCKS, c->d_small + 2048, c->stream)))
            goto done;
        if (hipMemcpyAsync(c->h_small, c->d_small + 2048, sizeof(double) * nt,
                           hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = -2; goto done; }
        double rmax = 0.0;
        std::vector<double> rall(nt);
        for (int r = 0; r < nt; r++) rall[r] = sqrt(c->h_small[r]);
        for (int r = 0; r < nroots; r++) { rn[r] = rall[r]; rmax = std::max(rmax, rn[r]); }
        if (p->verbose) {
            fprintf(stdout, " iter %3d  m %2d  %s  E0 %.12f  max|r| %.3e\n", it, m,
                    phase32 ? "fp32" : "fp64", theta[0] + p->totfrze, rmax);
            fflush(stdout);
        }
        if (phase32 && rmax < JCI_MIXED_GAIN * best32) { best32 = rmax; it_best32 = it; }
        if (phase32 && (rmax < p->mixed_switch || it == p->maxiter ||
                        (it - it_best32 >= JCI_MIXED_STALL &&
                         best32 < jci_mixed_near(p->mixed_switch)))) {
            // hand over to FP64 (below the switch residual, or stalled at the FP32 sigma's
            // noise floor above it): keep the lowest kc Ritz vectors of the FP32 phase (the
            // basis itself is FP64) and recompute their sigma in FP64
            phase32 = 0;
            int kc = std::min(m, krymin);
            if (kc < m) {
                std::vector<double> Yc((size_t)m * kc);
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < kc; j++) Yc[(size_t)i * kc + j] = Yf[(size_t)j * m + i];
                if ((rc = upload_small(w, Yc.data(), m * kc, &dY))) goto done;
                if ((rc = jk_rotate_inplace(m, kc, N, w.V, w.ld, dY, c->stream))) goto done;
                m = kc;
            }
            if ((rc = sigma_block(w, 0, m, cur_method()))) goto done;
            count_sigma(m);
            gvalid = 0;
            np = 0;
            continue;
        }
        bool req_done = true;
        std::vector<int> todo;
        for (int r = 0; r < nt; r++) {
            bool ok = rall[r] < p->restol;
            if (ok && r < nroots && p->etol > 0.0) {
                double gap = 1e300;
                for (int j = nroots; j < m; j++) gap = std::min(gap, fabs(theta[j] - theta[r]));
                ok = rall[r] * rall[r] < p->etol * gap;
            }
            if (!ok && r < nroots) req_done = false;
            if (r >= nroots && rall[r] < p->buftol) ok = true;
            if (!ok) todo.push_back(r);
        }
        if (req_done) { converged = 1; break; }
        if (it == p->maxiter) break;
        int k = (int)todo.size();
        // compact the unconverged corrections to the front of T
        for (int t = 0; t < k; t++)
            if (todo[t] != t &&
                hipMemcpyAsync(w.T + (size_t)t * w.ld, w.T + (size_t)todo[t] * w.ld, vb,
                               hipMemcpyDeviceToDevice, c->stream) != hipSuccess) { rc = -2; goto done; }
        if (m + k > krymax) {
            // collapse onto the lowest kc Ritz vectors (basis and sigma rotated together)
            int kc = std::max(krymin, nt);
            kc = std::min(kc, krymax - k);
            if (kc < nroots) kc = nroots;
            if (kc + k > krymax) k = krymax - kc;   // drop trailing buffer corrections
            if (p->restart == JCI_RESTART_GDK && np > 0) {
                // GD+k: also keep the previous iteration's Ritz vectors
                std::vector<double> C((size_t)m * (krymax - k)), Gc((size_t)(krymax - k) * (krymax - k));
                const int nc = jci_gdk_restart(m, G.data(), Yf.data(), kc, mp, np, Yp.data(),
                                               krymax - k, C.data(), Gc.data());
                if ((rc = upload_small(w, C.data(), m * nc, &dY))) goto done;
                if ((rc = jk_rotate_inplace(m, nc, N, w.V, w.ld, dY, c->stream))) goto done;
                if ((rc = jk_rotate_inplace(m, nc, N, w.S, w.ld, dY, c->stream))) goto done;
                m = nc;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) Gk[(size_t)i * JCI_MAXBASIS + j] = Gc[(size_t)i * m + j];
            } else {
                std::vector<double> Yc((size_t)m * kc);
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < kc; j++) Yc[(size_t)i * kc + j] = Yf[(size_t)j * m + i];
                if ((rc = upload_small(w, Yc.data(), m * kc, &dY))) goto done;
                if ((rc = jk_rotate_inplace(m, kc, N, w.V, w.ld, dY, c->stream))) goto done;
                if ((rc = jk_rotate_inplace(m, kc, N, w.S, w.ld, dY, c->stream))) goto done;
                m = kc;
                // the collapsed basis vectors are the Ritz vectors themselves: V^T S = diag(theta)
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) Gk[(size_t)i * JCI_MAXBASIS + j] = i == j ? theta[i] : 0.0;
            }
            Yf.assign((size_t)m * m, 0.0);
            for (int i = 0; i < m; i++) Yf[(size_t)i * m + i] = 1.0;
            gvalid = m;
            // this iteration's Ritz vectors are the leading kc columns of the new basis
            mp = m;
            np = std::min(nt, kc);
            Yp.assign((size_t)np * mp, 0.0);
            for (int j = 0; j < np; j++) Yp[(size_t)j * mp + j] = 1.0;
        } else {
            mp = m;
            np = nt;
            Yp.assign(Yf.begin(), Yf.begin() + (size_t)nt * m);
        }
        int kept = 0;
        if ((rc = orthonormalize_new(w, m, k, &kept))) goto done;
        if (kept == 0) break;   // no new directions: stagnated
        if (hipMemcpyAsync(w.V + (size_t)m * w.ld, w.T, vb * kept, hipMemcpyDeviceToDevice,
                           c->stream) != hipSuccess) { rc = -2; goto done; }
        if ((rc = sigma_block(w, m, kept, cur_method()))) goto done;
        count_sigma(kept);
        m += kept;
    }
    for (int r = 0; r < nroots; r++) evals[r] = theta[r] + p->totfrze;
    if (evecs) {
        // the lowest nv Ritz vectors, JCI_MAXK at a time through T; rows past the basis
        // size (nout > m) are zero
        const int nv = std::max(nroots, nout);
        for (int r0 = 0; r0 < nv; r0 += w.kb) {
            const int kb = std::min(w.kb, std::min(nv, m) - r0);
            if (kb <= 0) break;
            std::vector<double> Yc((size_t)m * kb);
            for (int i = 0; i < m; i++)
                for (int r = 0; r < kb; r++) Yc[(size_t)i * kb + r] = Yf[(size_t)(r0 + r) * m + i];
            if ((rc = upload_small(w, Yc.data(), m * kb, &dY))) goto done;
            if ((rc = jk_combine(m, kb, N, w.V, w.ld, dY, kb, w.T, w.ld, 0.0, c->stream))) goto done;
            for (int r = 0; r < kb; r++)
                if (hipMemcpyAsync(evecs + (size_t)(r0 + r) * N, w.T + (size_t)r * w.ld, sizeof(double) * N,
                                   hipMemcpyDeviceToHost, c->stream) != hipSuccess) { rc = -2; goto done; }
            if (hipStreamSynchronize(c->stream) != hipSuccess) { rc = -2; goto done; }
        }
        for (int r = std::min(nv, m); r < nv; r++) memset(evecs + (size_t)r * N, 0, sizeof(double) * N);
    }
    if (!converged && rc == 0) {
        jci_set_error("solve: not converged in %d iterations (max|r| above %.3e)", p->maxiter, p->restol);
        rc = 1;
    }
done:
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iterations = it;
        stats->converged = converged;
        stats->nbasis = m;
        stats->nsigma64 = w.ns64;
        stats->nsigma32 = w.ns32;
        for (int r = 0; r < nroots && r < JCI_MAXROOTS; r++) stats->rnorm[r] = rn[r];
        stats->sigma_seconds = w.sigma_time;
        stats->seconds = now_s() - t_start;
    }
    hipStreamSynchronize(c->stream);
    hipFree(w.V); hipFree(w.S); hipFree(w.T); hipFree(w.X); hipFree(w.Y);
    return rc;
}

static int ham_sigma(void *self, int method, int k, const double *x, double *y)
{
    return sigma_dev((jci_ham *)self, method, k, x, y);
}

// vector-major sigma; returns 1 when the method has no such form (the caller packs)
static int ham_sigma_cols(void *self, int method, int k, const double *V, long long ldv,
                          double *S, long long lds, double *stage)
{
    jci_ham *h = (jci_ham *)self;
    method = resolve_method(h, method, 0);
    if (method == JCI_SIGMA_FAST64 || method == JCI_SIGMA_FAST32) {
        // the table sigma reads determinant-major input and writes the columns directly
        RCHECK(jk_pack(k, h->sp->h.ndets, V, ldv, stage, h->sp->ctx->stream));
        return fast_run(h, method, k, stage, S, lds, nullptr);
    }
    if (!is_gf(method)) return 1;
    if (!h->have_gf) { jci_set_error("sigma: dense-string sigma not applicable"); return -1; }
    return gf_run(h, method == JCI_SIGMA_GEMM64 ? 64 : 32, k, V, ldv, S, lds, nullptr);
}

static int ham_masks(void *self, long long n, const long long *d_idx, uint64_t *oa, uint64_t *ob)
{
    jci_ham *h = (jci_ham *)self;
    return jk_det_masks(h->sp->d_desc, h->sp->d_mask[0], h->sp->d_mask[1], d_idx, 0, n, oa, ob,
                        h->sp->ctx->stream);
}

static int ham_resolve(void *self, int method, int want32)
{
    return resolve_method((jci_ham *)self, method, want32);
}

JCI_API int jci_solve(jci_ham *h, const jci_solve_params *p, const double *guess,
                         double *evals, double *evecs, jci_solve_stats *stats)
{
    SolveOp op{};
    op.self = h;
    op.ctx = h->sp->ctx;
    op.N = h->sp->h.ndets;
    op.d_hd = h->d_hd;
    op.d_m1 = h->d_m1;
    op.d_m2 = h->d_m2;
    op.sigma = ham_sigma;
    op.sigma_cols = ham_sigma_cols;
    op.masks = ham_masks;
    op.resolve = ham_resolve;
    return jci_solve_core(&op, p, guess, evals, evecs, stats, 0);
}

// the lowest max(nroots, nout) Ritz vectors into evecs (the MPI driver's restart file)
int jci_solve_nout(jci_ham *h, const jci_solve_params *p, const double *guess, double *evals,
                   double *evecs, jci_solve_stats *stats, int nout)
{
    SolveOp op{};
    op.self = h;
    op.ctx = h->sp->ctx;
    op.N = h->sp->h.ndets;
    op.d_hd = h->d_hd;
    op.d_m1 = h->d_m1;
    op.d_m2 = h->d_m2;
    op.sigma = ham_sigma;
    op.sigma_cols = ham_sigma_cols;
    op.masks = ham_masks;
    op.resolve = ham_resolve;
    return jci_solve_core(&op, p, guess, evals, evecs, stats, nout);
}

// the solver operator of a Hamiltonian with another preconditioner diagonal (device, N)
int jci_ham_solve_op(jci_ham *h, const double *d_pre, SolveOp *op)
{
    *op = SolveOp{};
    op->self = h;
    op->ctx = h->sp->ctx;
    op->N = h->sp->h.ndets;
    op->d_hd = d_pre ? d_pre : h->d_hd;
    op->d_m1 = h->d_m1;
    op->d_m2 = h->d_m2;
    op->sigma = ham_sigma;
    op->sigma_cols = ham_sigma_cols;
    op->masks = ham_masks;
    op->resolve = ham_resolve;
    return 0;
}

// ---- accessors for the batched engine ---------------------------------------------------------
int jci_ham_fastargs(jci_ham *h, FastArgs *out)
{
    if (!h->have_fast) { jci_set_error("batch: the Hamiltonian has no excitation tables"); return -1; }
    *out = h->fast[2];            // the batched sigma uses 256-string beta tiles
    return 0;
}
const double *jci_ham_hd(jci_ham *h) { return h->d_hd; }
long long jci_ham_ndets(jci_ham *h) { return h->sp->h.ndets; }
jci_ctx *jci_ham_ctx(jci_ham *h) { return h->sp->ctx; }

// the initial block of the Hamiltonian's next batched solve, computed ahead (any host thread,
// on the Hamiltonian's context) so that jci_batch_add only uploads and scatters it
JCI_API int jci_ham_prepare_guess(jci_ham *h, const jci_solve_params *p)
{
    JCI_CHECK(hipSetDevice(h->sp->ctx->device));
    SolveOp op{};
    op.self = h;
    op.ctx = h->sp->ctx;
    op.N = h->sp->h.ndets;
    op.d_hd = h->d_hd;
    op.d_m1 = h->d_m1;
    op.d_m2 = h->d_m2;
    op.masks = ham_masks;
    int ntrack, krymin, krymax;
    jci_subspace_sizes(p, op.N, &ntrack, &krymin, &krymax);
    jci_solve_params q = *p;
    q.guess = p->guess == JCI_GUESS_DIAG ? JCI_GUESS_DIAG : JCI_GUESS_BLOCK;
    std::vector<long long> idx;
    std::vector<double> U;
    RCHECK(guess_compute(&op, op.ctx, &q, krymin, idx, U));
    free(h->g_idx);
    free(h->g_U);
    h->g_idx = (long long *)malloc(sizeof(long long) * idx.size());
    h->g_U = (double *)malloc(sizeof(double) * U.size());
    if (!h->g_idx || !h->g_U) { jci_set_error("prepare_guess: out of host memory"); return -3; }
    memcpy(h->g_idx, idx.data(), sizeof(long long) * idx.size());
    memcpy(h->g_U, U.data(), sizeof(double) * U.size());
    h->g_nb = (int)idx.size();
    h->g_nguess = krymin;
    h->g_refdim = p->refdim;
    h->g_diag = q.guess == JCI_GUESS_DIAG;
    return 0;
}

// a prepared initial block for (refdim, nguess, diag): its size nb and host arrays, or 0
int jci_ham_prepared_guess(jci_ham *h, int refdim, int nguess, int diag_guess, int *nb,
                           const long long **idx, const double **U)
{
    if (!h->g_idx || h->g_refdim != refdim || h->g_nguess != nguess || h->g_diag != (diag_guess != 0))
        return 0;
    *nb = h->g_nb;
    *idx = h->g_idx;
    *U = h->g_U;
    return 1;
}

// the initial block applied to V on stream st; d_scr holds (8 + 8 nguess) nb bytes
int jci_guess_apply(int nb, const long long *idx, const double *U, int nguess, double *V,
                    long long ld, void *d_scr, hipStream_t st)
{
    std::vector<long long> vi(idx, idx + nb);
    std::vector<double> vu(U, U + (size_t)nb * nguess);
    return guess_apply(vi, vu, nguess, V, ld, d_scr, st);
}

// initial guess block of a Hamiltonian into V (nguess vectors, leading dimension ld)
int jci_ham_guess_block(jci_ham *h, int refdim, int nguess, int diag_guess, double *V, long long ld)
{
    SolveOp op{};
    op.self = h;
    op.ctx = h->sp->ctx;
    op.N = h->sp->h.ndets;
    op.d_hd = h->d_hd;
    op.d_m1 = h->d_m1;
    op.d_m2 = h->d_m2;
    op.sigma = ham_sigma;
    op.sigma_cols = ham_sigma_cols;
    op.masks = ham_masks;
    op.resolve = ham_resolve;
    Work w{};
    w.op = &op;
    w.c = op.ctx;
    w.N = op.N;
    w.ld = ld;
    w.V = V;
    jci_solve_params p;
    jci_solve_defaults(&p);
    p.refdim = refdim;
    p.guess = diag_guess ? JCI_GUESS_DIAG : JCI_GUESS_BLOCK;
    return guess_block(w, &p, nguess);
}

// ---- synthetic integrals ------------------------------------------------------------------
// SplitMix64 (Steele, Lea, Flood 2014) drives a uniform stream; normals by Box-Muller
// (cosine branch only, one normal per two uniforms).  Draw order: h_ij for i < j
// (row-major over 1-based i, j), then for P = 1..2n the upper triangle of L^P (row-major,
// diagonal first in each row).
struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal()
    {
        double u1 = 1.0 - uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

JCI_API int jci_synth_integrals(int n, uint64_t seed, double *m1, double *m2)
{
    if (n < 1 || n > 64) { jci_set_error("synth: norbs out of range"); return -1; }
    SplitMix rng{seed};
    size_t n1 = m1_len(n);
    std::vector<double> h((size_t)n * n, 0.0);
    for (int i = 1; i <= n; i++) h[(size_t)(i - 1) * n + (i - 1)] = -2.0 + 0.25 * i;
    for (int i = 1; i <= n; i++)
        for (int j = i + 1; j <= n; j++) {
            double v = 0.05 * rng.normal();
            h[(size_t)(i - 1) * n + (j - 1)] = h[(size_t)(j - 1) * n + (i - 1)] = v;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) m1[jci_tri(i, j)] = h[(size_t)i * n + j];
    // L[P][pair] with pair = tri(i, j)
    int np = 2 * n;
    std::vector<double> L((size_t)np * n1);
    for (int P = 0; P < np; P++)
        for (int i = 0; i < n; i++)
            for (int j = i; j < n; j++) {
                double v = (i == j) ? 0.1 + 0.2 * rng.uniform() : 0.02 * rng.normal();
                L[(size_t)P * n1 + jci_tri(i, j)] = v;
            }
    for (size_t a = 0; a < n1; a++)
        for (size_t b = 0; b <= a; b++) {
            double s = 0.0;
            for (int P = 0; P < np; P++) s += L[(size_t)P * n1 + a] * L[(size_t)P * n1 + b];
            m2[a * (a + 1) / 2 + b] = s;
        }
    return 0;
}

// dense-string sigma constants of a Hamiltonian (1 when it has none)
int jci_ham_gemmfci(jci_ham *h, GemmFci *out)
{
    if (!h->have_gf) return 1;
    *out = h->gf;
    return 0;
}
