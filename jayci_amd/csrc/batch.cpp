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
    void *mem = nullptr;               // the one device block V .. X32 are carved from
    size_t mem_bytes = 0;
    int m = 0, krymin = 0, krymax = 0, nroots = 0, ntrack = 0;
    int phase32 = 0;
    double best32 = 1e300;             // FP32 phase: smallest largest-residual so far, and when
    int it_best32 = 0;
    int sig_first = 0, sig_count = 0;  // basis columns whose sigma is pending
    int kc = 0;                        // collapse target of this step (0: none)
    int iters = 0;
    long long ns64 = 0, ns32 = 0;
    std::vector<double> theta, Yf;
    std::vector<double> G;             // projected matrix V^T S (m x m), kept across steps
    std::vector<double> Yp;            // previous step's Ritz vectors of the tracked pairs (GD+k
    int mp = 0, np = 0;                // restart): np vectors of mp rows, zero below mp
    std::vector<double> Cg;            // this step's GD+k restart block (m x kc, row-major)
    int gnew = 0;                      // first basis column whose sigma arrived this step
    double rn[JCI_MAXROOTS] = {0};
    double ev[JCI_MAXROOTS] = {0};
    std::vector<int> todo;
    int kept = 0;
    double t_add = 0, t_done = 0;
};

// pinned staging areas for descriptor / coefficient uploads, used in rotation: an area is
// rewritten only after the event recorded behind its last upload has completed, so the host
// never waits for the whole stream just to refill a buffer
#define JB_STAGES 4
struct Stage {
    BHeader *hH;
    BDesc *hD;
    double *hcoef;
    hipEvent_t ev;
    int used;
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
    Stage stg[JB_STAGES];
    int stg_cur;
    int desc_cap;      // descriptors per launch (dD and every stage's hD)
    BHeader *hHs;      // header of the current stage
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
    int bn64, bn32;    // E tile widths of the dense-string sigma
    long long steps;
    double sigma_seconds;
    // device blocks of released instances, kept for reuse (oldest first)
    std::vector<std::pair<void *, size_t>> cache;
    size_t cache_bytes;
    // device scratch of prepared initial blocks (grows only)
    void *gscr;
    size_t gscr_bytes;
    // host wall time per step phase: 0 sigma (all of run_sigma), 1 Gram + Rayleigh-Ritz,
    // 2 residuals and decisions, 3 selection / collapse / hand-over, 4 Gram-Schmidt and
    // orthonormalisation, 5 append
    double phase_seconds[6];
};

static int nblk_rows(long long N) { return (int)((N + 255) / 256); }
// reduction work-groups per instance: 4096 rows each (enough independent loads in flight to
// stream the vectors), at most 1024
// blocks of a reduction over N rows: `rows` per block, at most `cap` per instance.  The
// Gram kernel walks its rows 32 at a time and takes 1024 (a chain of 32 steps), at most 1024
// blocks (each leaves an m x w partial).  The residual kernel walks them 256 at a time and
// also takes 1024, at most 8192 blocks (k partials each), so that a 4M-row instance spreads
// over more blocks than the GPU holds at once.  The partition depends on N alone, so an
// instance's sums, and its result, do not depend on the other instances of a launch.
static int nblk_red(long long N, long long rows, long long cap)
{
    long long b = (N + rows - 1) / rows;
    return (int)std::max<long long>(1, std::min<long long>(b, cap));
}

// grow the partial-sum buffer (reductions) to hold `need` doubles
static int ensure_partial(jci_batch *b, size_t need)
{
    if (need <= b->partial_cap) return 0;
    JCI_CHECK(hipStreamSynchronize(b->c->stream));
    hipFree(b->dpartial);
    b->dpartial = nullptr;
    size_t cap = std::max(need, b->partial_cap * 2);
    if (hipMalloc(&b->dpartial, sizeof(double) * cap) != hipSuccess) {
        b->partial_cap = 0;
        jci_set_error("batch: out of device memory for %zu reduction partials", cap);
        return -3;
    }
    b->partial_cap = cap;
    return 0;
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
    // descriptors per launch: at most two per instance (V and S rotated together)
    b->desc_cap = 2 * capacity + 8;
    JCI_CHECK(hipMalloc(&b->dD, sizeof(BDesc) * b->desc_cap));
    JCI_CHECK(hipMalloc(&b->dS, sizeof(BSigma) * capacity * 2));
    JCI_CHECK(hipHostMalloc(&b->hS, sizeof(BSigma) * capacity * 2, 0));
    b->coef_cap = (size_t)capacity * (JCI_MAXBASIS * JCI_MAXBASIS + JCI_MAXK);
    JCI_CHECK(hipMalloc(&b->dcoef, sizeof(double) * b->coef_cap));
    for (int i = 0; i < JB_STAGES; i++) {
        Stage &g = b->stg[i];
        JCI_CHECK(hipHostMalloc(&g.hH, sizeof(BHeader) * (1 + JB_NCLASS), 0));
        JCI_CHECK(hipHostMalloc(&g.hD, sizeof(BDesc) * b->desc_cap, 0));
        JCI_CHECK(hipHostMalloc(&g.hcoef, sizeof(double) * b->coef_cap, 0));
        JCI_CHECK(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming));
        g.used = 0;
    }
    b->stg_cur = 0;
    b->hD = b->stg[0].hD;
    b->hcoef = b->stg[0].hcoef;
    b->hHs = b->stg[0].hH;
    b->out_cap = (size_t)capacity * JCI_MAXBASIS * JCI_MAXBASIS;
    JCI_CHECK(hipMalloc(&b->dout, sizeof(double) * b->out_cap));
    JCI_CHECK(hipHostMalloc(&b->hout, sizeof(double) * b->out_cap, 0));
    b->partial_cap = (size_t)1 << 22;
    JCI_CHECK(hipMalloc(&b->dpartial, sizeof(double) * b->partial_cap));
    JCI_CHECK(hipMalloc(&b->dGH, sizeof(GfHeader)));
    JCI_CHECK(hipHostMalloc(&b->hGH, sizeof(GfHeader), 0));
    JCI_CHECK(hipMalloc(&b->dGT, sizeof(GfTask) * capacity));
    JCI_CHECK(hipHostMalloc(&b->hGT, sizeof(GfTask) * capacity, 0));
    b->epool = nullptr;
    b->epool_bytes = 0;
    b->epool_want = (size_t)24 << 30;
    b->bn64 = GF_BN64_DEFAULT;
    b->bn32 = GF_BN32_DEFAULT;
    *out = b;
    return 0;
}

// Instance memory: one device block per instance, and released blocks are kept (up to
// JB_CACHE_BYTES, oldest freed first) for the next instance that fits in one at most twice
// its size.  A served stream admits and releases instances of a few GB each (V and S of a
// 4M-determinant request with a 32-vector subspace are 2 GB); fresh hipMalloc / hipFree
// of those took 15 and 1.3 ms per request, time in which no step runs.
#define JB_CACHE_BYTES (16ULL << 30)
static void cache_flush(jci_batch *b)
{
    for (auto &e : b->cache) hipFree(e.first);
    b->cache.clear();
    b->cache_bytes = 0;
}

static void *block_get(jci_batch *b, size_t bytes, size_t *got)
{
    int best = -1;
    for (size_t i = 0; i < b->cache.size(); i++) {
        const size_t sz = b->cache[i].second;
        if (sz >= bytes && sz / 2 <= bytes && (best < 0 || sz < b->cache[best].second)) best = (int)i;
    }
    if (best >= 0) {
        void *p = b->cache[best].first;
        *got = b->cache[best].second;
        b->cache_bytes -= *got;
        b->cache.erase(b->cache.begin() + best);
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        cache_flush(b);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *got = bytes;
    return p;
}

// the caller has synchronised the stream: no kernel still uses the block
static void free_inst(jci_batch *b, BInst &s)
{
    if (s.mem) {
        b->cache.push_back({s.mem, s.mem_bytes});
        b->cache_bytes += s.mem_bytes;
        while (b->cache_bytes > JB_CACHE_BYTES) {
            hipFree(b->cache.front().first);
            b->cache_bytes -= b->cache.front().second;
            b->cache.erase(b->cache.begin());
        }
    }
    s.mem = nullptr;
    s.mem_bytes = 0;
    s.V = s.S = s.T = s.X = s.Y = nullptr;
    s.X32 = nullptr;
}

JCI_API void jci_batch_destroy(jci_batch *b)
{
    if (!b) return;
    hipSetDevice(b->c->device);
    hipStreamSynchronize(b->c->stream);
    for (auto &s : b->in) if (s.used) free_inst(b, s);
    cache_flush(b);
    hipFree(b->gscr);
    for (auto &g : b->graphs) hipGraphExecDestroy(g.exec);
    hipFree(b->dH); hipHostFree(b->hH);
    hipFree(b->dD);
    for (int i = 0; i < JB_STAGES; i++) {
        hipHostFree(b->stg[i].hH); hipHostFree(b->stg[i].hD); hipHostFree(b->stg[i].hcoef);
        hipEventDestroy(b->stg[i].ev);
    }
    hipFree(b->dS); hipHostFree(b->hS);
    hipFree(b->dcoef);
    hipFree(b->dout); hipHostFree(b->hout);
    hipFree(b->dpartial);
    hipFree(b->dGH); hipHostFree(b->hGH);
    hipFree(b->dGT); hipHostFree(b->hGT);
    hipFree(b->epool);
    delete b;
}

JCI_API int jci_batch_set_graphs(jci_batch *b, int on) { b->use_graphs = on; return 0; }

JCI_API int jci_batch_set_tiles(jci_batch *b, int bn64, int bn32)
{
    if ((bn64 > 0 && !jk_gf_bn_valid(64, bn64)) || (bn32 > 0 && !jk_gf_bn_valid(32, bn32))) {
        jci_set_error("batch: tile widths %d / %d not available", bn64, bn32);
        return -1;
    }
    hipStreamSynchronize(b->c->stream);
    if (bn64 > 0) b->bn64 = bn64;
    if (bn32 > 0) b->bn32 = bn32;
    for (auto &g : b->graphs) hipGraphExecDestroy(g.exec);   // recorded with the old widths
    b->graphs.clear();
    return 0;
}

// E tile widths of the batch from timings of a representative full-CI Hamiltonian (the
// grouped launch runs every instance with one width per precision; the largest instances
// dominate its time): each available width timed with jci_bench_sigma, the fastest kept
JCI_API int jci_batch_tune_tiles(jci_batch *b, jci_ham *h, int nvec, int reps, int *bn64,
                                 int *bn32)
{
    const int c64[2] = {16, 32}, c32[2] = {32, 64};
    int best64 = 0, best32 = 0;
    double t64 = 1e300, t32 = 1e300;
    for (int i = 0; i < 2; i++) {
        int w64 = c64[i], w32 = c32[i];
        RCHECK(jci_ham_gemm_tiles(h, &w64, &w32));
        double ms = 0;
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_GEMM64, nvec, reps, &ms));
        if (ms < t64) { t64 = ms; best64 = c64[i]; }
        RCHECK(jci_bench_sigma(h, JCI_SIGMA_GEMM32, nvec, reps, &ms));
        if (ms < t32) { t32 = ms; best32 = c32[i]; }
    }
    RCHECK(jci_ham_gemm_tiles(h, &best64, &best32));
    RCHECK(jci_batch_set_tiles(b, best64, best32));
    if (bn64) *bn64 = best64;
    if (bn32) *bn32 = best32;
    return 0;
}

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
    // V, S: krymax vectors; T: JCI_MAXK; X, Y (table sigma only): JCI_MAXK; X32: JCI_MAXK FP32
    // (every piece a multiple of 512 bytes, as ld is a multiple of 64)
    const size_t need = vb * (2 * (size_t)s.krymax + (s.gf ? 1 : 3) * (size_t)JCI_MAXK) +
                        sizeof(float) * (size_t)s.ld * JCI_MAXK;
    s.mem = block_get(b, need, &s.mem_bytes);
    if (s.mem) {
        char *q = (char *)s.mem;
        s.V = (double *)q; q += vb * s.krymax;
        s.S = (double *)q; q += vb * s.krymax;
        s.T = (double *)q; q += vb * JCI_MAXK;
        if (!s.gf) {
            s.X = (double *)q; q += vb * JCI_MAXK;
            s.Y = (double *)q; q += vb * JCI_MAXK;
        }
        s.X32 = (float *)q;
    }
    if (!s.mem) {
        jci_set_error("batch: out of device memory for an instance of %lld determinants", s.N);
        return -3;
    }
    s.used = 1;
    s.t_add = tnow();
    // initial block: prepared ahead with the Hamiltonian (jci_ham_prepare_guess, any thread)
    // it is only uploaded and scattered here, on the batch stream
    int rc, gnb = 0;
    const long long *gidx = nullptr;
    const double *gU = nullptr;
    if (jci_ham_prepared_guess(h, p->refdim, s.krymin, p->guess == JCI_GUESS_DIAG, &gnb, &gidx, &gU)) {
        const size_t need = (sizeof(long long) + sizeof(double) * s.krymin) * (size_t)gnb;
        rc = 0;
        if (need > b->gscr_bytes) {
            hipFree(b->gscr);
            b->gscr = nullptr;
            b->gscr_bytes = 0;
            if (hipMalloc(&b->gscr, need) != hipSuccess) {
                (void)hipGetLastError();
                jci_set_error("batch: out of device memory for an initial block");
                rc = -3;
            } else {
                b->gscr_bytes = need;
            }
        }
        if (!rc) rc = jci_guess_apply(gnb, gidx, gU, s.krymin, s.V, s.ld, b->gscr, b->c->stream);
    } else {
        rc = jci_ham_guess_block(h, p->refdim, s.krymin, p->guess == JCI_GUESS_DIAG, s.V, s.ld);
    }
    if (rc) { free_inst(b, s); s.used = 0; return rc; }
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
    free_inst(b, b->in[slot]);
    b->in[slot] = BInst();
    return 0;
}

// ---- helpers ------------------------------------------------------------------------------
// move to the next staging area (waits only for that area's last uploads); b->hD, b->hcoef
// then point into it
static int stage_next(jci_batch *b)
{
    b->stg_cur = (b->stg_cur + 1) % JB_STAGES;
    Stage &g = b->stg[b->stg_cur];
    if (g.used) JCI_CHECK(hipEventSynchronize(g.ev));
    g.used = 0;
    b->hD = g.hD;
    b->hcoef = g.hcoef;
    b->hHs = g.hH;
    return 0;
}

// mark the current area busy until the uploads queued so far have run
static int stage_done(jci_batch *b)
{
    Stage &g = b->stg[b->stg_cur];
    JCI_CHECK(hipEventRecord(g.ev, b->c->stream));
    g.used = 1;
    return 0;
}

// hslot < 0: descriptors only (the launches read per-class headers, upload_class_headers)
static int upload_desc(jci_batch *b, int n, long long total, int hslot = 0)
{
    if (n > b->desc_cap) { jci_set_error("batch: %d descriptors above %d", n, b->desc_cap); return -1; }
    b->hHs->n = n;
    b->hHs->total = total;
    if (hslot >= 0)
        JCI_CHECK(hipMemcpyAsync(b->dH + hslot, b->hHs, sizeof(BHeader), hipMemcpyHostToDevice,
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

// Basis classes of the combine and Gram kernels: each class launches the instantiation
// sized for its largest basis (16, 32, 48 or 64 vectors), so small instances do not run with
// the registers and LDS of the largest one in the step.
static int basis_class(int m) { return m <= 16 ? 0 : m <= 32 ? 1 : m <= 48 ? 2 : 3; }
static const int kClassMax[JB_NCLASS] = {16, 32, 48, JCI_MAXBASIS};

// descriptor order grouped by class (stable inside a class); first[c] / cnt[c]: the class's
// range in that order
static void class_order(const std::vector<int> &cls, std::vector<int> &order, int first[JB_NCLASS],
                        int cnt[JB_NCLASS])
{
    order.clear();
    for (int c = 0; c < JB_NCLASS; c++) {
        first[c] = (int)order.size();
        for (int t = 0; t < (int)cls.size(); t++)
            if (cls[t] == c) order.push_back(t);
        cnt[c] = (int)order.size() - first[c];
    }
}

// one header per class into device slots JB_HCLASS.. (the descriptors go up with upload_desc)
static int upload_class_headers(jci_batch *b, const int cnt[JB_NCLASS], const long long tot[JB_NCLASS])
{
    for (int c = 0; c < JB_NCLASS; c++) {
        memset(&b->hHs[1 + c], 0, sizeof(BHeader));
        b->hHs[1 + c].n = cnt[c];
        b->hHs[1 + c].total = tot[c];
    }
    JCI_CHECK(hipMemcpyAsync(b->dH + JB_HCLASS, b->hHs + 1, sizeof(BHeader) * JB_NCLASS,
                             hipMemcpyHostToDevice, b->c->stream));
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
    int n = (int)ids.size();
    if (n == 0) { ofs.clear(); return 0; }
    if (n > b->desc_cap) { jci_set_error("batch: %d descriptors above %d", n, b->desc_cap); return -1; }
    RCHECK(stage_next(b));
    long long pofs = 0, tot[JB_NCLASS] = {0, 0, 0, 0};
    size_t o = 0;
    std::vector<int> cls(n), order;
    int first[JB_NCLASS], cnt[JB_NCLASS];
    ofs.assign(n, 0);
    for (int t = 0; t < n; t++) {
        if (m[t] < 1 || w[t] < 1 || m[t] > JCI_MAXBASIS || w[t] > JCI_MAXBASIS) {
            jci_set_error("batch: Gram of %d x %d outside 1..%d", m[t], w[t], JCI_MAXBASIS);
            return -1;
        }
        cls[t] = basis_class(std::max(m[t], w[t]));
        ofs[t] = o;
        o += (size_t)m[t] * w[t];
    }
    if (o > b->out_cap) { jci_set_error("batch: scratch too small"); return -1; }
    class_order(cls, order, first, cnt);
    for (int q = 0; q < n; q++) {
        const int t = order[q];
        BInst &s = b->in[ids[t]];
        BDesc &d = b->hD[q];
        memset(&d, 0, sizeof(d));
        d.N = s.N; d.ld = s.ld; d.A = A[t]; d.B = B[t]; d.m = m[t]; d.k = w[t];
        d.blk0 = tot[cls[t]]; d.nblk = nblk_red(s.N, 1024, 1024); d.pofs = pofs;
        d.out = b->dout + ofs[t];
        tot[cls[t]] += d.nblk;
        pofs += (long long)d.nblk * m[t] * w[t];
    }
    RCHECK(ensure_partial(b, (size_t)pofs));
    RCHECK(upload_desc(b, n, tot[0] + tot[1] + tot[2] + tot[3], -1));
    RCHECK(upload_class_headers(b, cnt, tot));
    RCHECK(stage_done(b));
    for (int c = 0; c < JB_NCLASS; c++)
        if (cnt[c] > 0)
            RCHECK(jb_gram(b->dH + JB_HCLASS + c, b->dD + first[c], cnt[c], kClassMax[c],
                           b->dpartial, b->c->stream));
    RCHECK(download_out(b, o));
    return 0;
}

// O_t = beta O_t + A_t C_t for the listed instances; the caller has moved to a fresh staging
// area (stage_next) and written the coefficients into hcoef at cofs
static int run_combine(jci_batch *b, const std::vector<int> &ids, const std::vector<double *> &A,
                       const std::vector<double *> &O, const std::vector<int> &m,
                       const std::vector<int> &k, const std::vector<size_t> &cofs,
                       double beta, size_t coef_len)
{
    int n = (int)ids.size();
    if (n == 0) return 0;
    if (n > b->desc_cap) { jci_set_error("batch: %d descriptors above %d", n, b->desc_cap); return -1; }
    long long tot[JB_NCLASS] = {0, 0, 0, 0};
    std::vector<int> cls(n), order;
    int first[JB_NCLASS], cnt[JB_NCLASS];
    for (int t = 0; t < n; t++) {
        if (m[t] > JCI_MAXBASIS || k[t] > m[t]) {
            jci_set_error("batch: combine of %d vectors into %d (basis limit %d)", m[t], k[t],
                          JCI_MAXBASIS);
            return -1;
        }
        cls[t] = basis_class(m[t]);
    }
    class_order(cls, order, first, cnt);
    for (int q = 0; q < n; q++) {
        const int t = order[q];
        BInst &s = b->in[ids[t]];
        BDesc &d = b->hD[q];
        memset(&d, 0, sizeof(d));
        d.N = s.N; d.ld = s.ld; d.A = A[t]; d.O = O[t]; d.m = m[t]; d.k = k[t];
        d.C = b->dcoef + cofs[t]; d.beta = beta;
        d.blk0 = tot[cls[t]]; d.nblk = nblk_rows(s.N);
        tot[cls[t]] += d.nblk;
    }
    RCHECK(upload_coef(b, coef_len));
    RCHECK(upload_desc(b, n, tot[0] + tot[1] + tot[2] + tot[3], -1));
    RCHECK(upload_class_headers(b, cnt, tot));
    RCHECK(stage_done(b));
    for (int c = 0; c < JB_NCLASS; c++)
        if (cnt[c] > 0 && jb_combine(b->dH + JB_HCLASS + c, b->dD + first[c], kClassMax[c],
                                     b->c->stream))
            return -1;
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
    for (int i : ids)   // one vector of the largest instance, with the row padding
        need = std::max(need, (size_t)b->in[i].g.La * (size_t)(b->in[i].N + GF_EALIGN) * esz);
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
            // E of k vectors: La rows of jk_gf_lde(k N) elements
            const size_t per = (size_t)s.g.La * s.N * esz;
            const size_t pad = (size_t)s.g.La * GF_EALIGN * esz;
            const size_t avail = b->epool_bytes - used;
            const size_t room = avail > pad ? (avail - pad) / per : 0;
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
            t.lde = jk_gf_lde((long long)k * s.N);
            long long u[4];
            jk_gf_units(prec, prec == 64 ? b->bn64 : b->bn32, t, u);
            for (int q = 0; q < 4; q++) { t.base[q] = base[q]; base[q] += u[q]; }
            used += ((size_t)s.g.La * (size_t)t.lde * esz + 255) / 256 * 256;
            who.push_back(i);
            kk.push_back(k);
            n++;
        }
        if (n == 0) {
            for (int i : ids)
                if (b->in[i].sig_count > 0) {
                    jci_set_error("batch: sigma pool too small for instance %d", i);
                    return -3;
                }
            break;
        }
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
                int rc = jk_gf_group(prec, prec == 64 ? b->bn64 : b->bn32, b->dGH, b->dGT, nullptr,
                                     c->stream);
                hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc) return rc;
                if (e != hipSuccess) { jci_set_error("batch: graph capture failed: %s", hipGetErrorString(e)); return -2; }
                JCI_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                hipGraphDestroy(g);
                b->graphs.push_back(GraphKey{prec, -1, ex});
            }
            JCI_CHECK(hipGraphLaunch(ex, c->stream));
        } else {
            RCHECK(jk_gf_group(prec, prec == 64 ? b->bn64 : b->bn32, b->dGH, b->dGT, b->hGH,
                               c->stream));
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
// Per step: (1) sigma of every pending basis column (grouped); (2) the new columns of the
// projected matrix G = V^T S (one batched Gram; the rest of G is kept) and a host eigensolve
// per instance; (3) residuals of the tracked Ritz pairs with the diagonal preconditioner;
// (4) decisions (precision hand-over, convergence, collapse); (5) collapse / hand-over
// rotations; (6) the selected corrections appended to V; (7) one Gram [V_old | V_new]^T V_new
// gives both the projection C = V_old^T V_new and the block's own overlap; the projected
// overlap is V_new^T V_new - C^T C, and one combine writes (V_new - V_old C) L in place, L its
// inverse square root on the kept directions.  Where the projection cancels most of a
// vector, an explicit projection pass runs first (Gram-Schmidt twice).
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
    double tp = tnow();
    // 1. pending sigma (FP32 and FP64 groups)
    std::vector<int> g32, g64;
    for (int i : act) {
        BInst &s = b->in[i];
        s.gnew = s.sig_count > 0 ? s.sig_first : s.m;
        if (s.sig_count > 0) (s.phase32 ? g32 : g64).push_back(i);
    }
    RCHECK(run_sigma(b, g32, 32));
    RCHECK(run_sigma(b, g64, 64));
    b->phase_seconds[0] += tnow() - tp;
    tp = tnow();
    // 2. new columns of G and Rayleigh-Ritz
    {
        std::vector<int> ids;
        std::vector<double *> A, B;
        std::vector<int> m, w;
        std::vector<size_t> ofs;
        for (int i : act) {
            BInst &s = b->in[i];
            if (s.G.size() != (size_t)s.gnew * s.gnew) s.gnew = 0;   // no valid leading block
            if (s.gnew >= s.m) continue;
            ids.push_back(i); A.push_back(s.V); B.push_back(s.S + (size_t)s.gnew * s.ld);
            m.push_back(s.m); w.push_back(s.m - s.gnew);
        }
        RCHECK(run_gram(b, ids, A, B, m, w, ofs));
        for (size_t t = 0; t < ids.size(); t++) {
            BInst &s = b->in[ids[t]];
            const int mm = s.m, g0 = s.gnew, ww = mm - g0;
            std::vector<double> G((size_t)mm * mm, 0.0);
            for (int i = 0; i < g0; i++)
                for (int j = 0; j < g0; j++) G[(size_t)i * mm + j] = s.G[(size_t)i * g0 + j];
            const double *o = b->hout + ofs[t];
            for (int i = 0; i < mm; i++)
                for (int j = g0; j < mm; j++) {
                    double v = o[(size_t)i * ww + (j - g0)];
                    if (i >= g0) v = 0.5 * (v + o[(size_t)j * ww + (i - g0)]);
                    G[(size_t)i * mm + j] = G[(size_t)j * mm + i] = v;
                }
            s.G.swap(G);
        }
    }
    for (int i : act) {
        BInst &s = b->in[i];
        s.theta.assign(s.m, 0.0);
        s.Yf.assign((size_t)s.m * s.m, 0.0);
        jci_jacobi_eig(s.m, s.G.data(), s.theta.data(), s.Yf.data());
        s.iters++;
    }
    b->phase_seconds[1] += tnow() - tp;
    tp = tnow();
    // 3. residuals / corrections -> T, norms
    std::vector<size_t> ofs(act.size());
    {
        RCHECK(stage_next(b));
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
            d.out = b->dout + o; d.blk0 = blk; d.nblk = nblk_red(s.N, 1024, 8192); d.pofs = pofs;
            ofs[t] = o;
            co += (size_t)m * k + k;
            o += k;
            blk += d.nblk;
            pofs += (long long)d.nblk * k;
        }
        RCHECK(upload_coef(b, co));
        RCHECK(upload_desc(b, (int)act.size(), blk));
        RCHECK(stage_done(b));
        RCHECK(ensure_partial(b, (size_t)pofs));
        RCHECK(jb_residual(b->dH, b->dD, (int)act.size(), b->dpartial, c->stream));
        RCHECK(download_out(b, o));
    }
    // 4. decisions per instance
    std::vector<int> grow;          // instances adding vectors this step
    std::vector<int> handover;      // FP32 -> FP64: collapse onto Ritz vectors, re-multiply
    std::vector<int> collapse;      // full subspace: collapse onto kc Ritz vectors
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
        if (s.phase32 && rmax < JCI_MIXED_GAIN * s.best32) { s.best32 = rmax; s.it_best32 = s.iters; }
        if (s.phase32 && (rmax < s.p.mixed_switch || s.iters >= s.p.maxiter ||
                          (s.iters - s.it_best32 >= JCI_MIXED_STALL &&
                           s.best32 < jci_mixed_near(s.p.mixed_switch)))) {
            s.phase32 = 0;
            handover.push_back(act[t]);
            continue;
        }
        bool req_done = true;
        s.todo.clear();
        for (int r = 0; r < nt; r++) {
            bool ok = rall[r] < s.p.restol;
            if (ok && r < s.nroots && s.p.etol > 0.0) {
                double gap = 1e300;
                for (int j = s.nroots; j < s.m; j++) gap = std::min(gap, fabs(s.theta[j] - s.theta[r]));
                ok = rall[r] * rall[r] < s.p.etol * gap;
            }
            if (!ok && r < s.nroots) req_done = false;
            if (r >= s.nroots && rall[r] < s.p.buftol) ok = true;
            if (!ok) s.todo.push_back(r);
        }
        if (req_done) { s.done = 1; s.t_done = tnow(); continue; }
        if (s.iters >= s.p.maxiter) { s.done = 1; s.failed = 1; s.t_done = tnow(); continue; }
        // a full subspace collapses onto kc Ritz vectors; buffer corrections that do not fit
        // next to them are dropped
        s.kc = 0;
        if (s.m + (int)s.todo.size() > s.krymax) {
            int kc = std::max(s.krymin, nt);
            kc = std::min(kc, s.krymax - (int)s.todo.size());
            if (kc < s.nroots) kc = s.nroots;
            if (kc + (int)s.todo.size() > s.krymax) s.todo.resize(s.krymax - kc);
            s.kc = kc;
            s.Cg.clear();
            if (s.p.restart == JCI_RESTART_GDK && s.np > 0) {
                // GD+k: the lowest kc Ritz vectors and the previous step's ones
                const int kmax = s.krymax - (int)s.todo.size();
                std::vector<double> Gc((size_t)kmax * kmax);
                s.Cg.resize((size_t)s.m * kmax);
                s.kc = jci_gdk_restart(s.m, s.G.data(), s.Yf.data(), kc, s.mp, s.np, s.Yp.data(),
                                       kmax, s.Cg.data(), Gc.data());
                s.Cg.resize((size_t)s.m * s.kc);
                Gc.resize((size_t)s.kc * s.kc);
                s.Yp.swap(Gc);   // holds C^T G C until the rotation is issued
            }
            // this step's Ritz vectors of the tracked pairs: after the collapse they are
            // the leading columns of the basis
            s.np = std::min(nt, kc);
            collapse.push_back(act[t]);
        } else {
            s.mp = s.m;
            s.np = nt;
            s.Yp.assign(s.Yf.begin(), s.Yf.begin() + (size_t)nt * s.m);
        }
        grow.push_back(act[t]);
    }
    b->phase_seconds[2] += tnow() - tp;
    tp = tnow();
    // 5. rotations: hand-over (V onto its lowest kc Ritz vectors; their FP64 sigma is pending
    // for the next step) and collapse (V and S onto kc Ritz vectors, G becomes diagonal)
    {
        RCHECK(stage_next(b));
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        auto add_rot = [&](BInst &s, int i, int kc, bool with_s) {
            if (!s.Cg.empty() && with_s)
                memcpy(b->hcoef + co, s.Cg.data(), sizeof(double) * (size_t)s.m * kc);
            else
                for (int r = 0; r < s.m; r++)
                    for (int j = 0; j < kc; j++) b->hcoef[co + (size_t)r * kc + j] = s.Yf[(size_t)j * s.m + r];
            ids.push_back(i); A.push_back(s.V); O.push_back(s.V); m.push_back(s.m); k.push_back(kc);
            cofs.push_back(co);
            if (with_s) {
                ids.push_back(i); A.push_back(s.S); O.push_back(s.S); m.push_back(s.m); k.push_back(kc);
                cofs.push_back(co);
            }
            co += (size_t)s.m * kc;
        };
        for (int i : handover) {
            BInst &s = b->in[i];
            add_rot(s, i, std::min(s.m, s.krymin), false);
        }
        for (int i : collapse) {
            BInst &s = b->in[i];
            add_rot(s, i, s.kc, true);
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
        for (int i : handover) {
            BInst &s = b->in[i];
            s.m = std::min(s.m, s.krymin);
            s.G.clear();
            s.sig_first = 0;
            s.sig_count = s.m;
            s.np = 0;
        }
        for (int i : collapse) {
            BInst &s = b->in[i];
            const int kc = s.kc;
            if (!s.Cg.empty()) {
                s.G.swap(s.Yp);   // C^T G C
                s.Cg.clear();
            } else {
                s.G.assign((size_t)kc * kc, 0.0);
                for (int j = 0; j < kc; j++) s.G[(size_t)j * kc + j] = s.theta[j];
            }
            s.m = kc;
            s.mp = kc;
            s.Yp.assign((size_t)s.np * s.mp, 0.0);
            for (int j = 0; j < s.np; j++) s.Yp[(size_t)j * s.mp + j] = 1.0;
        }
    }
    if (grow.empty()) { b->phase_seconds[3] += tnow() - tp; return 0; }
    // 6. append the selected corrections: V[m .. m+nk) = T Sel
    {
        RCHECK(stage_next(b));
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        for (int i : grow) {
            BInst &s = b->in[i];
            const int nt = std::min(s.ntrack, (int)s.theta.size());
            const int nk = (int)s.todo.size();
            for (int r = 0; r < nt; r++)
                for (int j = 0; j < nk; j++) b->hcoef[co + (size_t)r * nk + j] = (s.todo[j] == r) ? 1.0 : 0.0;
            ids.push_back(i); A.push_back(s.T); O.push_back(s.V + (size_t)s.m * s.ld); m.push_back(nt);
            k.push_back(nk); cofs.push_back(co);
            co += (size_t)nt * nk;
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
    }
    b->phase_seconds[3] += tnow() - tp;
    tp = tnow();
    // 7. orthonormalise the new blocks against the basis and among themselves
    std::vector<int> pend = grow, fin;
    std::vector<std::vector<double>> W(b->cap);   // final coefficients per slot
    for (int pass = 0; pass < 2 && !pend.empty(); pass++) {
        std::vector<double *> A, B;
        std::vector<int> m, w;
        std::vector<size_t> gofs;
        for (int i : pend) {
            BInst &s = b->in[i];
            const int nk = (int)s.todo.size();
            A.push_back(s.V); B.push_back(s.V + (size_t)s.m * s.ld);
            m.push_back(s.m + nk); w.push_back(nk);
        }
        RCHECK(run_gram(b, pend, A, B, m, w, gofs));
        std::vector<int> redo;
        for (size_t t = 0; t < pend.size(); t++) {
            BInst &s = b->in[pend[t]];
            const int mo = s.m, nk = (int)s.todo.size();
            const double *g = b->hout + gofs[t];           // (mo + nk) x nk
            // C = V_old^T V_new (mo x nk), TT = V_new^T V_new, TTp = TT - C^T C
            std::vector<double> TT((size_t)nk * nk), TTp((size_t)nk * nk);
            for (int i = 0; i < nk; i++)
                for (int j = 0; j < nk; j++) {
                    double v = 0.5 * (g[(size_t)(mo + i) * nk + j] + g[(size_t)(mo + j) * nk + i]);
                    double cc = 0.0;
                    for (int r = 0; r < mo; r++) cc += g[(size_t)r * nk + i] * g[(size_t)r * nk + j];
                    TT[(size_t)i * nk + j] = v;
                    TTp[(size_t)i * nk + j] = v - cc;
                }
            std::vector<double> ev(nk), U((size_t)nk * nk);
            jci_jacobi_eig(nk, TTp.data(), ev.data(), U.data());
            double emax = 0.0, cancel = 1.0;
            for (double e : ev) emax = std::max(emax, e);
            for (int j = 0; j < nk; j++)
                if (TT[(size_t)j * nk + j] > 0) cancel = std::min(cancel, TTp[(size_t)j * nk + j] / TT[(size_t)j * nk + j]);
            const bool weak = cancel < 0.5 || ev[0] < 1e-6 * emax;
            if (pass == 0 && weak) { redo.push_back(pend[t]); continue; }
            // kept directions (largest first) and W = [-C L; L]
            int nkk = 0;
            for (int i = nk - 1; i >= 0; i--) if (ev[i] > 1e-12 * emax && ev[i] > 1e-24) nkk++;
            s.kept = nkk;
            if (nkk == 0) { fin.push_back(pend[t]); continue; }
            std::vector<double> L((size_t)nk * nkk);
            int col = 0;
            for (int i = nk - 1; i >= 0 && col < nkk; i--) {
                if (!(ev[i] > 1e-12 * emax && ev[i] > 1e-24)) continue;
                const double sc = 1.0 / sqrt(ev[i]);
                for (int r = 0; r < nk; r++) L[(size_t)r * nkk + col] = U[(size_t)i * nk + r] * sc;
                col++;
            }
            std::vector<double> &Wt = W[pend[t]];
            Wt.assign((size_t)(mo + nk) * nkk, 0.0);
            for (int r = 0; r < mo; r++)
                for (int j = 0; j < nkk; j++) {
                    double v = 0.0;
                    for (int q = 0; q < nk; q++) v += g[(size_t)r * nk + q] * L[(size_t)q * nkk + j];
                    Wt[(size_t)r * nkk + j] = -v;
                }
            for (int r = 0; r < nk; r++)
                for (int j = 0; j < nkk; j++) Wt[(size_t)(mo + r) * nkk + j] = L[(size_t)r * nkk + j];
            fin.push_back(pend[t]);
        }
        if (!redo.empty()) {
            // explicit projection V_new -= V_old C, then the Gram again
            RCHECK(stage_next(b));
            std::vector<double *> AA, O;
            std::vector<int> mm, kk;
            std::vector<size_t> cofs;
            size_t co = 0;
            for (size_t t = 0; t < pend.size(); t++) {
                if (std::find(redo.begin(), redo.end(), pend[t]) == redo.end()) continue;
                BInst &s = b->in[pend[t]];
                const int mo = s.m, nk = (int)s.todo.size();
                const double *g = b->hout + gofs[t];
                for (int e = 0; e < mo * nk; e++) b->hcoef[co + e] = -g[e];
                AA.push_back(s.V); O.push_back(s.V + (size_t)mo * s.ld); mm.push_back(mo); kk.push_back(nk);
                cofs.push_back(co);
                co += (size_t)mo * nk;
            }
            RCHECK(run_combine(b, redo, AA, O, mm, kk, cofs, 1.0, co));
        }
        pend.swap(redo);
    }
    // final combine: V[m .. m+kept) = V[0 .. m+nk) W (in place: each row is read before it is
    // written)
    {
        RCHECK(stage_next(b));
        std::vector<int> ids;
        std::vector<double *> A, O;
        std::vector<int> m, k;
        std::vector<size_t> cofs;
        size_t co = 0;
        for (int i : fin) {
            BInst &s = b->in[i];
            if (s.kept == 0) { s.done = 1; s.failed = 1; s.t_done = tnow(); continue; }
            const int nk = (int)s.todo.size();
            const std::vector<double> &Wt = W[i];
            memcpy(b->hcoef + co, Wt.data(), sizeof(double) * Wt.size());
            ids.push_back(i); A.push_back(s.V); O.push_back(s.V + (size_t)s.m * s.ld);
            m.push_back(s.m + nk); k.push_back(s.kept); cofs.push_back(co);
            co += Wt.size();
        }
        RCHECK(run_combine(b, ids, A, O, m, k, cofs, 0.0, co));
        for (int i : ids) {
            BInst &s = b->in[i];
            s.sig_first = s.m;
            s.sig_count = s.kept;
            s.m += s.kept;
        }
    }
    b->phase_seconds[4] += tnow() - tp;
    return 0;
}

JCI_API int jci_batch_finished(jci_batch *b, int *slots, int cap)
{
    int n = 0;
    for (int i = 0; i < b->cap; i++)
        if (b->in[i].used && b->in[i].done) {
            if (slots && n < cap) slots[n] = i;
            n++;
        }
    return n;
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

JCI_API int jci_batch_phase_seconds(const jci_batch *b, double *out6)
{
    for (int i = 0; i < 6; i++) out6[i] = b->phase_seconds[i];
    return 0;
}
