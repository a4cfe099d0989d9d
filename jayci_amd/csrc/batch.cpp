// Batched block Davidson over many CI instances (one launch per step for all of them).
//
// Every active instance advances one Davidson iteration per jci_batch_step: a single
// batched sigma for all new vectors (persistent kernel, recorded once per (precision, K)
// as a HIP graph and replayed), a batched MFMA Gram, host Rayleigh-Ritz per instance,
// a batched residual/preconditioner, batched two-pass Gram-Schmidt and symmetric
// orthonormalisation, and a batched append.  Instances may be added between steps (the
// serving loop admits requests this way) and leave when converged.  Each instance keeps
// its own precision phase: FP32 sigma until its residual norm falls below mixed_switch,
// then its basis collapses onto the current Ritz vectors, which are re-multiplied in FP64.
// Full-CI instances take the dense-string MFMA sigma (grouped kernels over all of them, the
// opposite-spin intermediates in a shared pool, several rounds when the pool is full); the
// others take the persistent table kernel.
#include "batch.h"
#include "slater.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <algorithm>
#include <vector>

#define RCHECK(x)                      \
    do {                               \
        int rc_ = (x);                 \
        if (rc_) return rc_;           \
    } while (0)

static double tnow()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void jci_jacobi_eig(int n, const double *a, double *eval, double *vec);   // jci.cpp

struct BInst {
    int used = 0, done = 0, failed = 0;
    jci_ham *h = nullptr;
    jci_solve_params p;
    FastArgs fa;
    int gf = 0;                        // dense-string sigma
    GemmFci g;
    long long N = 0, ld = 0;
    double *V = nullptr, *S = nullptr, *T = nullptr, *X = nullptr, *Y = nullptr;
    float *X32 = nullptr;
    int m = 0, krymin = 0, krymax = 0, nroots = 0, ntrack = 0;
    int phase32 = 0;
    int sig_first = 0, sig_count = 0;  // basis columns whose sigma is pending
    int iters = 0;
    long long ns64 = 0, ns32 = 0;
    std::vector<double> theta, Yf;
    double rn[JCI_MAXROOTS] = {0};
    double ev[JCI_MAXROOTS] = {0};
    std::vector<int> todo;
    int kept = 0;
    double t_add = 0, t_done = 0;
};

struct GraphKey {
    int prec, K;
    hipGraphExec_t exec;
};

struct jci_batch {
    jci_ctx *c;
    int cap;
    std::vector<BInst> in;
    int W;          // tile width of the batched sigma
    int m1max;      // largest orbital-pair count (LDS sizing)
    // device / pinned arenas
    BHeader *dH, *hH;
    BDesc *dD, *hD;
    BSigma *dS, *hS;
    double *dcoef, *hcoef;
    size_t coef_cap;
    double *dout, *hout;
    size_t out_cap;
    double *dpartial;
    size_t partial_cap;
    std::vector<GraphKey> graphs;
    int use_graphs;
    // dense-string sigma: launch descriptors and the pool of opposite-spin intermediates
    GfHeader *dGH, *hGH;
    GfTask *dGT, *hGT;
    void *epool;
    size_t epool_bytes, epool_want;
    long long steps;
    double sigma_seconds;
};

static int nblk_rows(long long N) { return (int)((N + 255) / 256); }
static int nblk_red(long long N)
{
    long long b = (N + 8191) / 8192;
    return (int)std::max<long long>(1, std::min<long long>(b, 64));
}

JCI_API int jci_batch_create(jci_ctx *c, int capacity, int max_norbs, jci_batch **out)
{
    *out = nullptr;
    if (capacity < 1 || capacity > 4096) { jci_set_error("batch: capacity 1..4096"); return -1; }
    JCI_CHECK(hipSetDevice(c->device));
    jci_batch *b = new jci_batch();
    b->c = c;
    b->cap = capacity;
    b->in.resize(capacity);
    b->W = 256;
    b->m1max = max_norbs * (max_norbs + 1) / 2;
    b->use_graphs = 1;
    JCI_CHECK(hipMalloc(&b->dH, sizeof(BHeader) * 8));
    JCI_CHECK(hipHostMalloc(&b->hH, sizeof(BHeader) * 8, 0));
    JCI_CHECK(hipMalloc(&b->dD, sizeof(BDesc) * capacity));
    JCI_CHECK(hipHostMalloc(&b->hD, sizeof(BDesc) * capacity, 0));
    JCI_CHECK(hipMalloc(&b->dS, sizeof(BSigma) * capacity * 2));
    JCI_CHECK(hipHostMalloc(&b->hS, sizeof(BSigma) * capacity * 2, 0));
    b->coef_cap = (size_t)capacity * (JCI_MAXBASIS * JCI_MAXBASIS + JCI_MAXK);
    JCI_CHECK(hipMalloc(&b->dcoef, sizeof(double) * b->coef_cap));
    JCI_CHECK(hipHostMalloc(&b->hcoef, sizeof(double) * b->coef_cap, 0));
    b->out_cap = (size_t)capacity * JCI_MAXBASIS * JCI_MAXBASIS;
    JCI_CHECK(hipMalloc(&b->dout, sizeof(double) * b->out_cap));
    JCI_CHECK(hipHostMalloc(&b->hout, sizeof(double) * b->out_cap, 0));
    b->partial_cap = (size_t)capacity * 64 * JCI_MAXBASIS * JCI_MAXBASIS;
    JCI_CHECK(hipMalloc(&b->dpartial, sizeof(double) * b->partial_cap));
    JCI_CHECK(hipMalloc(&b->dGH, sizeof(GfHeader)));
    JCI_CHECK(hipHostMalloc(&b->hGH, sizeof(GfHeader), 0));
    JCI_CHECK(hipMalloc(&b->dGT, sizeof(GfTask) * capacity));
    JCI_CHECK(hipHostMalloc(&b->hGT, sizeof(GfTask) * capacity, 0));
    b->epool = nullptr;
    b->epool_bytes = 0;
    b->epool_want = (size_t)24 << 30;
    *out = b;
    return 0;
}

static void free_inst(BInst &s)
{
    hipFree(s.V); hipFree(s.S); hipFree(s.T); hipFree(s.X); hipFree(s.Y); hipFree(s.X32);
    s.V = s.S = s.T = s.X = s.Y = nullptr;
    s.X32 = nullptr;
}

JCI_API void jci_batch_destroy(jci_batch *b)
{
    if (!b) return;
    hipSetDevice(b->c->device);
    hipStreamSynchronize(b->c->stream);
    for (auto &s : b->in) if (s.used) free_inst(s);
    for (auto &g : b->graphs) hipGraphExecDestroy(g.exec);
    hipFree(b->dH); hipHostFree(b->hH);
    hipFree(b->dD); hipHostFree(b->hD);
    hipFree(b->dS); hipHostFree(b->hS);
    hipFree(b->dcoef); hipHostFree(b->hcoef);
    hipFree(b->dout); hipHostFree(b->hout);
    hipFree(b->dpartial);
    hipFree(b->dGH); hipHostFree(b->hGH);
    hipFree(b->dGT); hipHostFree(b->hGT);
    hipFree(b->epool);
    delete b;
}

JCI_API int jci_batch_set_graphs(jci_batch *b, int on) { b->use_graphs = on; return 0; }

JCI_API int jci_batch_set_pool(jci_batch *b, long long bytes)
{
    if (bytes < (1LL << 20)) { jci_set_error("batch: pool below 1 MiB"); return -1; }
    hipStreamSynchronize(b->c->stream);
    hipFree(b->epool);
    b->epool = nullptr;
    b->epool_bytes = 0;
    b->epool_want = (size_t)bytes;
    return 0;
}

// bytes of device memory an instance of N determinants takes in the batch
JCI_API long long jci_batch_bytes(long long N, int krymax)
{
    long long ld = (N + 63) / 64 * 64;
    return (long long)sizeof(double) * ld * (2LL * krymax + 3LL * JCI_MAXK) + (long long)sizeof(float) * ld * JCI_MAXK;
}

JCI_API int jci_batch_add(jci_batch *b, jci_ham *h, const jci_solve_params *p, int *slot)
{
    int k = -1;
    for (int i = 0; i < b->cap; i++) if (!b->in[i].used) { k = i; break; }
    if (k < 0) { jci_set_error("batch: no free slot"); return -1; }
    FastArgs fa;
    RCHECK(jci_ham_fastargs(h, &fa));
    if (fa.W != b->W) { jci_set_error("batch: tile width of the Hamiltonian differs"); return -1; }
    if (fa.m1 > b->m1max) { jci_set_error("batch: more orbitals than the batch was created for"); return -1; }
    BInst &s = b->in[k];
    s = BInst();
    s.h = h;
    s.p = *p;
    s.fa = fa;
    s.gf = jci_ham_gemmfci(h, &s.g) == 0;
    s.N = jci_ham_ndets(h);
    s.ld = (s.N + 63) / 64 * 64;
    s.nroots = p->nroots;
    s.krymax = (int)std::min<long long>(p->krymax, s.N);
    if (s.nroots < 1 || s.nroots > JCI_MAXROOTS || s.nroots > s.N || s.krymax > JCI_MAXBASIS ||
        p->nbuffer < 0) {
        jci_set_error("batch: bad root / subspace sizes");
        return -1;
    }
    s.ntrack = (int)std::min(std::min<long long>(s.nroots + p->nbuffer, JCI_MAXROOTS),
                             std::min<long long>(s.N, std::max(s.nroots, s.krymax / 2)));
    s.krymin = std::min(std::max(p->krymin, s.ntrack), s.krymax);
    if (s.krymin < s.nroots) {
        jci_set_error("batch: bad root / subspace sizes");
        return -1;
    }
    size_t vb = sizeof(double) * (size_t)s.ld;
    if (hipMalloc(&s.V, vb * s.krymax) != hipSuccess || hipMalloc(&s.S, vb * s.krymax) != hipSuccess ||
        hipMalloc(&s.T, vb * JCI_MAXK) != hipSuccess ||
        (!s.gf && (hipMalloc(&s.X, vb * JCI_MAXK) != hipSuccess ||
                   hipMalloc(&s.Y, vb * JCI_MAXK) != hipSuccess)) ||
        hipMalloc(&s.X32, sizeof(float) * s.ld * JCI_MAXK) != hipSuccess) {
        free_inst(s);
        jci_set_error("batch: out of device memory for an instance of %lld determinants", s.N);
        return -3;
    }
    s.used = 1;
    s.t_add = tnow();
    int rc = jci_ham_guess_block(h, p->refdim, s.krymin, p->guess == JCI_GUESS_DIAG, s.V, s.ld);
    if (rc) { free_inst(s); s.used = 0; return rc; }
    s.m = s.krymin;
    s.sig_first = 0;
    s.sig_count = s.krymin;
    s.phase32 = p->precision == JCI_PREC_MIXED;
    *slot = k;
    return 0;
}

JCI_API int jci_batch_release(jci_batch *b, int slot)
{
    if (slot < 0 || slot >= b->cap || !b->in[slot].used) { jci_set_error("batch: bad slot"); return -1; }
    hipStreamSynchronize(b->c->stream);
    free_inst(b->in[slot]);
    b->in[slot] = BInst();
    return 0;
}

// ---- helpers ------------------------------------------------------------------------------
static int upload_desc(jci_batch *b, int n, long long total, int hslot = 0)
{
    b->hH[hslot].n = n;
    b->hH[hslot].total = total;
    JCI_CHECK(hipMemcpyAsync(b->dH + hslot, b->hH + hslot, sizeof(BHeader), hipMemcpyHostToDevice,
                             b->c->stream));
    if (n > 0)
        JCI_CHECK(hipMemcpyAsync(b->dD, b->hD, sizeof(BDesc) * n, hipMemcpyHostToDevice, b->c->stream));
    return 0;
}

static int upload_coef(jci_batch *b, size_t len)
{
    if (len) JCI_CHECK(hipMemcpyAsync(b->dcoef, b->hcoef, sizeof(double) * len, hipMemcpyHostToDevice,
                                      b->c->stream));
    return 0;
}

static int download_out(jci_batch *b, size_t len)
{
    if (len) JCI_CHECK(hipMemcpyAsync(b->hout, b->dout, sizeof(double) * len, hipMemcpyDeviceToHost,
                                      b->c->stream));
    JCI_CHECK(hipStreamSynchronize(b->c->stream));
    return 0;
}

// Gram out_i = A_i^T B_i for the listed instances; result in hout at ofs[i]
static int run_gram(jci_batch *b, const std::vector<int> &ids, std::vector<double *> A,
                    std::vector<double *> B, std::vector<int> m, std::vector<int> w,
                    std::vector<size_t> &ofs)
{
    // the pinned staging buffers are rewritten below: earlier copies must have drained
    JCI_CHECK(hipStreamSynchronize(b->c->stream));
    int n = (int)ids.size();
    long long blk = 0, pofs = 0;
    size_t o = 0;
    ofs.assign(n, 0);
    for (int t = 0; t < n; t++) {
        BInst &s = b->in[ids[t]];
        BDesc &d = b->hD[t];
        memset(&d, 0, sizeof(d));
        d.N = s.N; d.ld = s.ld; d.A = A[t]; d.B = B[t]; d.m = m[t]; d.k = w[t];
        d.blk0 = blk; d.nblk = nblk_red(s.N); d.pofs = pofs;
        d.out = b->dout + o;
        ofs[t] = o;
        blk += d.nblk;
        pofs += (long long)d.nblk * m[t] * w[t];
        o += (size_t)m[t] * w[t];
    }
    if ((size_t)pofs > b->partial_cap || o > b->out_cap) { jci_set_error("batch: scratch too small"); return -1; }
    RCHECK(upload_desc(b, n, blk));
    RCHECK(jb_gram(b->dH, b->dD, n, b->dpartial, b->c->stream));
    RCHECK(download_out(b, o));
    return 0;
}

// O_t = beta O_t + A_t C_t for the listed instances (coefficients already in hcoef at cofs)
static int run_combine(jci_batch *b, const std::vector<int> &ids, const std::vector<double *> &A,
                       const std::vector<double *> &O, const std::vector<int> &m,
                       const std::vector<int> &k, const std::vector<size_t> &cofs,
                       double beta, size_t coef_len)
{
    int n = (int)ids.size();
    if (n == 0) return 0;
    JCI_CHECK(hipStreamSynchronize(b->c->stream));
    long long blk = 0;
    for (int t = 0; t < n; t++) {
        BInst &s = b->in[ids[t]];
        BDesc &d = b->hD[t];
        memset(&d, 0, sizeof(d));
        d.N = s.N; d.ld = s.ld; d.A = A[t]; d.O = O[t]; d.m = m[t]; d.k = k[t];
        d.C = b->dcoef + cofs[t]; d.beta = beta;
        d.blk0 = blk; d.nblk = nblk_rows(s.N);
        blk += d.nblk;
    }
    RCHECK(upload_coef(b, coef_len));
    RCHECK(upload_desc(b, n, blk));
    RCHECK(jb_combine(b->dH, b->dD, b->c->stream));
    return 0;
}

static hipGraphExec_t find_graph(jci_batch *b, int prec, int K);
static int run_sigma_tab(jci_batch *b, const std::vector<int> &ids, int prec);

// dense-string sigma of every pending column of the listed instances, in rounds that fit the
// intermediate pool
static int run_sigma_gf(jci_batch *b, const std::vector<int> &ids, int prec)
{
    if (ids.empty()) return 0;
    jci_ctx *c = b->c;
    const size_t esz = prec == 64 ? 8 : 4;
    size_t need = 0;
    for (int i : ids) need = std::max(need, (size_t)b->in[i].g.m1 * (size_t)b->in[i].N * esz);
    if (!b->epool || b->epool_bytes < need) {
        JCI_CHECK(hipStreamSynchronize(c->stream));
        hipFree(b->epool);
        b->epool = nullptr;
        b->epool_bytes = 0;
        size_t want = std::max(need, b->epool_want);
        if (hipMalloc(&b->epool, want) != hipSuccess) {
            jci_set_error("batch: out of device memory for the sigma pool (%zu bytes)", want);
            return -3;
        }
        b->epool_bytes = want;
    }
    for (;;) {
        JCI_CHECK(hipStreamSynchronize(c->stream));   // pinned descriptors free again
        int n = 0;
        size_t used = 0;
        long long base[4] = {0, 0, 0, 0};
        std::vector<int> who, kk;
        for (int i : ids) {
            BInst &s = b->in[i];
            if (s.sig_count <= 0) continue;
            const size_t per = (size_t)s.g.m1 * s.N * esz;
            const size_t room = (b->epool_bytes - used) / per;
            const int k = (int)std::min<size_t>(std::min<size_t>(s.sig_count, JCI_MAXK), room);
            if (k < 1) continue;
            GfTask &t = b->hGT[n];
            memset(&t, 0, sizeof(t));
            t.g = s.g;
            t.K = k;
            t.x64 = s.V + (size_t)s.sig_first * s.ld;
            t.ldx64 = s.ld;
            t.x32 = prec == 32 ? s.X32 : nullptr;
            t.y = s.S + (size_t)s.sig_first * s.ld;
            t.ldy = s.ld;
            t.E = (char *)b->epool + used;
            long long u[4];
            jk_gf_units(prec, t, u);
            for (int q = 0; q < 4; q++) { t.base[q] = base[q]; base[q] += u[q]; }
            used += (per * k + 255) / 256 * 256;
            who.push_back(i);
            kk.push_back(k);
            n++;
        }
        if (n == 0) break;
        b->hGH->n = n;
        for (int q = 0; q < 4; q++) b->hGH->total[q] = base[q];
        JCI_CHECK(hipMemcpyAsync(b->dGH, b->hGH, sizeof(GfHeader), hipMemcpyHostToDevice, c->stream));
        JCI_CHECK(hipMemcpyAsync(b->dGT, b->hGT, sizeof(GfTask) * n, hipMemcpyHostToDevice, c->stream));
        double t0 = tnow();
        if (b->use_graphs) {
            hipGraphExec_t ex = find_graph(b, prec, -1);
            if (!ex) {
                hipGraph_t g;
                JCI_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int rc = jk_gf_group(prec, b->dGH, b->dGT, nullptr, c->stream);
                hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc) return rc;
                if (e != hipSuccess) { jci_set_error("batch: graph capture failed: %s", hipGetErrorString(e)); return -2; }
                JCI_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                hipGraphDestroy(g);
                b->graphs.push_back(GraphKey{prec, -1, ex});
            }
            JCI_CHECK(hipGraphLaunch(ex, c->stream));
        } else {
            RCHECK(jk_gf_group(prec, b->dGH, b->dGT, b->hGH, c->stream));
        }
        JCI_CHECK(hipStreamSynchronize(c->stream));
        b->sigma_seconds += tnow() - t0;
        for (size_t t = 0; t < who.size(); t++) {
            BInst &s = b->in[who[t]];
            s.sig_first += kk[t];
            s.sig_count -= kk[t];
            if (prec == 64) s.ns64 += kk[t]; else s.ns32 += kk[t];
        }
    }
    return 0;
}

// sigma of every pending column of the listed instances of one precision
static int run_sigma(jci_batch *b, const std::vector<int> &ids, int prec)
{
    std::vector<int> gfi, tab;
    for (int i : ids) (b->in[i].gf ? gfi : tab).push_back(i);
    RCHECK(run_sigma_gf(b, gfi, prec));
    // the table kernel takes at most JCI_MAXK columns per instance and launch
    for (;;) {
        std::vector<int> now;
        std::vector<int> rest;
        for (int i : tab) {
            BInst &s = b->in[i];
            if (s.sig_count <= 0) continue;
            now.push_back(i);
            rest.push_back(s.sig_count - std::min(s.sig_count, JCI_MAXK));
            s.sig_count = std::min(s.sig_count, JCI_MAXK);
        }
        if (now.empty()) break;
        std::vector<int> first(now.size());
        for (size_t t = 0; t < now.size(); t++) {
            first[t] = b->in[now[t]].sig_first + b->in[now[t]].sig_count;
        }
        RCHECK(run_sigma_tab(b, now, prec));
        for (size_t t = 0; t < now.size(); t++) {
            BInst &s = b->in[now[t]];
            s.sig_first = first[t];
            s.sig_count = rest[t];
        }
    }
    return 0;
}

static hipGraphExec_t find_graph(jci_batch *b, int prec, int K)
{
    for (auto &g : b->graphs) if (g.prec == prec && g.K == K) return g.exec;
    return nullptr;
}

// table-kernel sigma for up to JCI_MAXK pending columns of every listed instance (the
// caller has cut sig_count to at most JCI_MAXK)
static int run_sigma_tab(jci_batch *b, const std::vector<int> &ids, int prec)
{
    int n = (int)ids.size();
    if (n == 0) return 0;
    jci_ctx *c = b->c;
    JCI_CHECK(hipStreamSynchronize(c->stream));
    // pack V columns [sig_first, +cnt) into X (determinant-major), optionally FP32 copy
    long long blk = 0;
    int K = 1;
    for (int t = 0; t < n; t++) {
        BInst &s = b->in[ids[t]];
        BDesc &d = b->hD[t];
        memset(&d, 0, sizeof(d));
        d.N = s.N; d.ld = s.ld; d.A = s.V; d.O = s.X; d.m = s.sig_first; d.k = s.sig_count;
        d.blk0 = blk; d.nblk = nblk_rows(s.N);
        blk += d.nblk;
        K = std::max(K, s.sig_count);
    }
    if (K > JCI_MAXK) { jci_set_error("batch: more than %d pending sigma vectors", JCI_MAXK); return -1; }
    RCHECK(upload_desc(b, n, blk, 0));
    RCHECK(jb_pack(b->dH, b->dD, 1, c->stream));
    // FP32 copy descriptors (A = X, O = X32) in header slot 1, descriptors after the pack's
    long long ntile = 0;
    for (int t = 0; t < n; t++) {
        BInst &s = b->in[ids[t]];
        BSigma &q = b->hS[t];
        q.a = s.fa;
        q.x = prec == 64 ? (const void *)s.X : (const void *)s.X32;
        q.y = s.Y;
        q.k = s.sig_count;
        q.tile0 = ntile;
        ntile += s.fa.ntiles;
    }
    JCI_CHECK(hipStreamSynchronize(c->stream));      // pack descriptors consumed
    if (prec == 32) {
        long long blk2 = 0;
        for (int t = 0; t < n; t++) {
            BInst &s = b->in[ids[t]];
            BDesc &d = b->hD[t];
            memset(&d, 0, sizeof(d));
            d.N = s.N; d.k = s.sig_count; d.A = s.X; d.O = (double *)s.X32;
            d.blk0 = blk2; d.nblk = (int)((s.N * s.sig_count + 255) / 256);
            blk2 += d.nblk;
        }
        RCHECK(upload_desc(b, n, blk2, 0));
        RCHECK(jb_to32(b->dH, b->dD, c->stream));
    }
    b->hH[1].n = n;
    b->hH[1].total = ntile;
    b->hH[1].kmax = K;
    JCI_CHECK(hipMemcpyAsync(b->dH + 1, b->hH + 1, sizeof(BHeader), hipMemcpyHostToDevice, c->stream));
    JCI_CHECK(hipMemcpyAsync(b->dS, b->hS, sizeof(BSigma) * n, hipMemcpyHostToDevice, c->stream));
    int Kt = K <= 1 ? 1 : (K <= 2 ? 2 : (K <= 4 ? 4 : 8));
    double t0 = tnow();
    if (b->use_graphs) {
        hipGraphExec_t ex = find_graph(b, prec, Kt);
        if (!ex) {
            // record the persistent sigma launch once for this (precision, K)
            hipGraph_t g;
            JCI_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = jb_sigma(prec, b->W, Kt, b->dH + 1, b->dS, b->m1max, c->stream);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) { jci_set_error("batch: graph capture failed: %s", hipGetErrorString(e)); return -2; }
            JCI_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
            hipGraphDestroy(g);
            b->graphs.push_back(GraphKey{prec, Kt, ex});
        }
        JCI_CHECK(hipGraphLaunch(ex, c->stream));
    } else {
        RCHECK(jb_sigma(prec, b->W, Kt, b->dH + 1, b->dS, b->m1max, c->stream));
    }
    JCI_CHECK(hipStreamSynchronize(c->stream));
    b->sigma_seconds += tnow() - t0;
    // unpack Y into S columns
    blk = 0;
    for (int t = 0; t < n; t++) {
        BInst &s = b->in[ids[t]];
        BDesc &d = b->hD[t];
        memset(&d, 0, sizeof(d));
        d.N = s.N; d.ld = s.ld; d.A = s.Y; d.O = s.S; d.m = s.sig_first; d.k = s.sig_count;
        d.blk0 = blk; d.nblk = nblk_rows(s.N);
        blk += d.nblk;
        if (prec == 64) s.ns64 += s.sig_count; else s.ns32 += s.sig_count;
        s.sig_count = 0;
    }
    RCHECK(upload_desc(b, n, blk, 0));
    RCHECK(jb_pack(b->dH, b->dD, 0, c->stream));
    JCI_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- one lockstep iteration -----------------------------------------------------------------
JCI_API int jci_batch_step(jci_batch *b, int *nactive)
{
    jci_ctx *c = b->c;
    JCI_CHECK(hipSetDevice(c->device));
    std::vector<int> act;
    for (int i = 0; i < b->cap; i++)
        if (b->in[i].used && !b->in[i].done) act.push_back(i);
    *nactive = (int)act.size();
    if (act.empty()) return 0;
    b->steps++;
    // 1. pending sigma (FP32 and FP64 groups)
    std::vector<int> g32, g64;
    for (int i : act) {
        BInst &s = b->in[i];
        if (s.sig_count > 0) (s.phase32 ? g32 : g64).push_back(i);
    }
    RCHECK(run_sigma(b, g32, 32));
    RCHECK(run_sigma(b, g64, 64));
    // 2. Gram V^T S and Rayleigh-Ritz
    std::vector<size_t> ofs;
    {
        std::vector<double *> A, B;
        std::vector<int> m, w;
        for (int i : act) {
            BInst &s = b->in[i];
            A.push_back(s.V); B.push_back(s.S); m.push_back(s.m); w.push_back(s.m);
        }
        RCHECK(run_gram(b, act, A, B, m, w, ofs));
    }
    for (size_t t = 0; t < act.size(); t++) {
        BInst &s = b->in[act[t]];
        int m = s.m;
        std::vector<double> G(b->hout + ofs[t], b->hout + ofs[t] + (size_t)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < i; j++) G[i * m + j] = G[j * m + i] = 0.5 * (G[i * m + j] + G[j * m + i]);
        s.theta.assign(m, 0.0);
        s.Yf.assign((size_t)m * m, 0.0);
        jci_jacobi_eig(m, G.data(), s.theta.data(), s.Yf.data());
        s.iters++;
    }
    // 3. residuals / corrections -> T, norms
    JCI_CHECK(hipStreamSynchronize(c->stream));
    {
        size_t co = 0, o = 0;
        long long blk = 0, pofs = 0;
        for (size_t t = 0; t < act.size(); t++) {
            BInst &s = b->in[act[t]];
            int m = s.m, k = std::min(s.ntrack, s.m);
            BDesc &d = b->hD[t];
            memset(&d, 0, sizeof(d));
            for (int i = 0; i < m; i++)
                for (int r = 0; r < k; r++) b->hcoef[co + (size_t)i * k + r] = s.Yf[(size_t)r * m + i];
            for (int r = 0; r < k; r++) b->hcoef[co + (size_t)m * k + r] = s.theta[r];
            d.N = s.N; d.ld = s.ld; d.A = s.V; d.B = s.S; d.O = s.T; d.m = m; d.k = k;
            d.C = b->dcoef + co; d.th = b->dcoef + co + (size_t)m * k; d.hd = s.fa.hd;
            d.out = b->dout + o; d.blk0 = blk; d.nblk = nblk_red(s.N); d.pofs = pofs;
            ofs[t] = o;
            co += (size_t)m * k + k;
            o += k;
            blk += d.nblk;
            pofs += (long long)d.nblk * k;
        }
        RCHECK(upload_coef(b, co));
        RCHECK(upload_desc(b, (int)act.size(), blk));
        RCHECK(jb_residual(b->dH, b->dD, (int)act.size(), b->dpartial, c->stream));
        RCHECK(download_out(b, o));
    }
    // 4. decisions per instance
    std::vector<int> grow;          // instances adding vectors this step
    std::vector<int> handover;      // FP32 -> FP64: collapse onto Ritz vectors, re-multiply
    for (size_t t = 0; t < act.size(); t++) {
        BInst &s = b->in[act[t]];
        const int nt = std::min(s.ntrack, s.m);
        double rall[JCI_MAXK];
        for (int r = 0; r < nt; r++) rall[r] = sqrt(std::max(0.0, b->hout[ofs[t] + r]));
        double rmax = 0.0;
        for (int r = 0; r < s.nroots; r++) {
            s.rn[r] = rall[r];
            rmax = std::max(rmax, s.rn[r]);
        }
        for (int r = 0; r < s.nroots; r++) s.ev[r] = s.theta[r] + s.p.totfrze;
        if (s.phase32 && (rmax < s.p.mixed_switch || s.iters >= s.p.maxiter)) {
            s.phase32 = 0;
            handover.push_back(act[t]);
            continue;
        }
        bool req_done = true;
        for (int r = 0; r < s.nroots; r++)
            if (!(s.rn[r] < s.p.restol)) req_done = false;
        if (req_done) { s.done = 1; s.t_done = tnow(); continue; }
        if (s.iters >= s.p.maxiter) { s.done = 1; s.failed = 1; s.t_done = tnow(); continue; }
        s.todo.clear();
        for (int r = 0; r < nt; r++)
            if (!(rall[r] < s.p.restol)) s.todo.push_back(r);
        // a full subspace collapses onto kc Ritz vectors; buffer corrections that do not fit
        // next to them are dropped
        if (s.m + (int)s.todo.size() > s.krymax) {
            int kc = std::max(s.krymin, nt);
            kc = std::min(kc, s.krymax - (int)s.todo.size());
            if (kc < s.nroots) kc = s.nroots;
            if (kc + (int)s.todo.size() > s.krymax) s.todo.resize(s.krymax - kc);
        }
        grow.push_back(act[t]);
    }
    // hand-over: V <- V Y[:, :kc] (the lowest kc Ritz vectors of the FP32 phase); their FP64
    // sigma is pending for the next step
    if (!handover.empty()) {
        JCI_CHECK(hipStreamSynchronize(c->stream));
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        for (int i : handover) {
            BInst &s = b->in[i];
            int kc = std::min(s.m, s.krymin);
            for (int r = 0; r < s.m; r++)
                for (int j = 0; j < kc; j++) b->hcoef[co + (size_t)r * kc + j] = s.Yf[(size_t)j * s.m + r];
            ids.push_back(i); A.push_back(s.V); O.push_back(s.V); m.push_back(s.m); k.push_back(kc);
            cofs.push_back(co);
            co += (size_t)s.m * kc;
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
        for (size_t t = 0; t < ids.size(); t++) {
            BInst &s = b->in[ids[t]];
            s.m = k[t];
            s.sig_first = 0;
            s.sig_count = s.m;
        }
    }
    if (grow.empty()) return 0;
    // 5. select unconverged corrections (in place) and collapse full subspaces
    JCI_CHECK(hipStreamSynchronize(c->stream));
    {
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        for (int i : grow) {
            BInst &s = b->in[i];
            int nk = (int)s.todo.size();
            const int nt = std::min(s.ntrack, s.m);
            bool ident = nk == nt;
            for (int j = 0; j < nk && ident; j++) ident = s.todo[j] == j;
            if (ident) continue;
            for (int r = 0; r < nt; r++)
                for (int j = 0; j < nk; j++) b->hcoef[co + (size_t)r * nk + j] = (s.todo[j] == r) ? 1.0 : 0.0;
            ids.push_back(i); A.push_back(s.T); O.push_back(s.T); m.push_back(nt); k.push_back(nk);
            cofs.push_back(co);
            co += (size_t)nt * nk;
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
    }
    JCI_CHECK(hipStreamSynchronize(c->stream));
    {
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        std::vector<int> newm;
        for (int i : grow) {
            BInst &s = b->in[i];
            int nk = (int)s.todo.size();
            if (s.m + nk <= s.krymax) continue;
            int kc = std::max(s.krymin, std::min(s.ntrack, s.m));
            kc = std::min(kc, s.krymax - nk);
            if (kc < s.nroots) kc = s.nroots;
            for (int r = 0; r < s.m; r++)
                for (int j = 0; j < kc; j++) b->hcoef[co + (size_t)r * kc + j] = s.Yf[(size_t)j * s.m + r];
            // rotate V and S with the same coefficients
            ids.push_back(i); A.push_back(s.V); O.push_back(s.V); m.push_back(s.m); k.push_back(kc); cofs.push_back(co);
            ids.push_back(i); A.push_back(s.S); O.push_back(s.S); m.push_back(s.m); k.push_back(kc); cofs.push_back(co);
            co += (size_t)s.m * kc;
            newm.push_back(kc);
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
        for (size_t t = 0, u = 0; t < ids.size(); t += 2, u++) b->in[ids[t]].m = newm[u];
    }
    // 6. two classical Gram-Schmidt passes of T against V
    for (int pass = 0; pass < 2; pass++) {
        std::vector<int> ids;
        std::vector<double *> A, B;
        std::vector<int> m, w;
        for (int i : grow) {
            BInst &s = b->in[i];
            ids.push_back(i); A.push_back(s.V); B.push_back(s.T); m.push_back(s.m);
            w.push_back((int)s.todo.size());
        }
        RCHECK(run_gram(b, ids, A, B, m, w, ofs));
        std::vector<double *> AA, O;
        std::vector<int> mm, kk;
        std::vector<size_t> cofs;
        size_t co = 0;
        JCI_CHECK(hipStreamSynchronize(c->stream));
        for (size_t t = 0; t < ids.size(); t++) {
            BInst &s = b->in[ids[t]];
            int nk = (int)s.todo.size();
            for (int e = 0; e < s.m * nk; e++) b->hcoef[co + e] = -b->hout[ofs[t] + e];
            AA.push_back(s.V); O.push_back(s.T); mm.push_back(s.m); kk.push_back(nk); cofs.push_back(co);
            co += (size_t)s.m * nk;
        }
        RCHECK(run_combine(b, ids, AA, O, mm, kk, cofs, 1.0, co));
    }
    // 7. symmetric orthonormalisation inside T
    {
        std::vector<double *> A, B;
        std::vector<int> m, w;
        for (int i : grow) {
            BInst &s = b->in[i];
            int nk = (int)s.todo.size();
            A.push_back(s.T); B.push_back(s.T); m.push_back(nk); w.push_back(nk);
        }
        RCHECK(run_gram(b, grow, A, B, m, w, ofs));
        std::vector<int> ids;
        std::vector<double *> AA, O;
        std::vector<int> mm, kk;
        std::vector<size_t> cofs;
        size_t co = 0;
        for (size_t t = 0; t < grow.size(); t++) {
            BInst &s = b->in[grow[t]];
            int k = (int)s.todo.size();
            std::vector<double> g(b->hout + ofs[t], b->hout + ofs[t] + (size_t)k * k), ev(k), U((size_t)k * k);
            for (int i = 0; i < k; i++)
                for (int j = 0; j < i; j++) g[i * k + j] = g[j * k + i] = 0.5 * (g[i * k + j] + g[j * k + i]);
            jci_jacobi_eig(k, g.data(), ev.data(), U.data());
            double emax = 0.0;
            for (double e : ev) emax = std::max(emax, e);
            int nk = 0;
            for (int i = k - 1; i >= 0; i--) if (ev[i] > 1e-12 * emax && ev[i] > 1e-24) nk++;
            s.kept = nk;
            if (nk == 0) continue;
            int col = 0;
            for (int i = k - 1; i >= 0 && col < nk; i--) {
                if (!(ev[i] > 1e-12 * emax && ev[i] > 1e-24)) continue;
                double sc = 1.0 / sqrt(ev[i]);
                for (int r = 0; r < k; r++) b->hcoef[co + (size_t)r * nk + col] = U[(size_t)i * k + r] * sc;
                col++;
            }
            ids.push_back(grow[t]); AA.push_back(s.T); O.push_back(s.T); mm.push_back(k); kk.push_back(nk);
            cofs.push_back(co);
            co += (size_t)k * nk;
        }
        RCHECK(run_combine(b, ids, AA, O, mm, kk, cofs, 0.0, co));
    }
    // 8. append T to V: V[m .. m+kept) = T (identity coefficients)
    JCI_CHECK(hipStreamSynchronize(c->stream));
    {
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        for (int i : grow) {
            BInst &s = b->in[i];
            if (s.kept == 0) { s.done = 1; s.failed = 1; s.t_done = tnow(); continue; }
            for (int r = 0; r < s.kept; r++)
                for (int j = 0; j < s.kept; j++) b->hcoef[co + (size_t)r * s.kept + j] = (r == j) ? 1.0 : 0.0;
            ids.push_back(i); A.push_back(s.T); O.push_back(s.V + (size_t)s.m * s.ld); m.push_back(s.kept);
            k.push_back(s.kept); cofs.push_back(co);
            co += (size_t)s.kept * s.kept;
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
        for (int i : ids) {
            BInst &s = b->in[i];
            s.sig_first = s.m;
            s.sig_count = s.kept;
            s.m += s.kept;
        }
    }
    JCI_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

JCI_API int jci_batch_result(jci_batch *b, int slot, int *done, double *evals, double *evecs,
                             jci_solve_stats *st)
{
    if (slot < 0 || slot >= b->cap || !b->in[slot].used) { jci_set_error("batch: bad slot"); return -1; }
    BInst &s = b->in[slot];
    *done = s.done ? (s.failed ? 2 : 1) : 0;
    if (evals) for (int r = 0; r < s.nroots; r++) evals[r] = s.ev[r];
    if (st) {
        memset(st, 0, sizeof(*st));
        st->iterations = s.iters;
        st->converged = s.done && !s.failed;
        st->nbasis = s.m;
        st->nsigma64 = s.ns64;
        st->nsigma32 = s.ns32;
        for (int r = 0; r < s.nroots; r++) st->rnorm[r] = s.rn[r];
        st->seconds = (s.done ? s.t_done : tnow()) - s.t_add;
    }
    if (evecs && s.done) {
        JCI_CHECK(hipStreamSynchronize(b->c->stream));
        // Ritz vectors: T = V Y (first nroots)
        std::vector<int> ids{slot};
        std::vector<double *> A{s.V}, O{s.T};
        std::vector<int> m{s.m}, k{s.nroots};
        std::vector<size_t> cofs{0};
        for (int i = 0; i < s.m; i++)
            for (int r = 0; r < s.nroots; r++) b->hcoef[(size_t)i * s.nroots + r] = s.Yf[(size_t)r * s.m + i];
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, (size_t)s.m * s.nroots));
        for (int r = 0; r < s.nroots; r++)
            JCI_CHECK(hipMemcpyAsync(evecs + (size_t)r * s.N, s.T + (size_t)r * s.ld, sizeof(double) * s.N,
                                     hipMemcpyDeviceToHost, b->c->stream));
        JCI_CHECK(hipStreamSynchronize(b->c->stream));
    }
    return 0;
}

JCI_API int jci_batch_run(jci_batch *b, int max_steps)
{
    for (int t = 0; t < max_steps; t++) {
        int na = 0;
        RCHECK(jci_batch_step(b, &na));
        if (na == 0) return 0;
    }
    return 0;
}

JCI_API double jci_batch_sigma_seconds(const jci_batch *b) { return b->sigma_seconds; }
