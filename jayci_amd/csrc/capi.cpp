// Reference-named drop-in entry points (C ABI) on top of the HIP runtime.
//
//   citrunc_rtnlist ... source/citruncate.c:123      generate_cimap ... source/cimapping.c:92
//   deallocate_cimap .. source/cimapping.c:70        compute_hdgls .... source/action_util.c:106
//   hmatels ........... source/action_util.c:204     compute_hv ....... source/action_util.c:121
//   dvdalg ............ source/davidson.c:91         print_civectors .. source/ioutil.c:184
//   read_civectors .... source/ioutil.c:215          write_wavefunction source/write_wavefunction.c:12
//   readgeninput ...... source/ioutil.c:332          readdaiinput ..... source/ioutil.c:280
//
// These operate on the reference's explicit determinant list (struct det).  The list is
// turned into per-determinant alpha/beta masks on the device; generate_cimap runs the
// O(N^2) excitation-degree screen on the GPU, returns the reference's section lists and
// keeps the connectivity on the device, keyed by the hmap pointer, so that compute_hv and
// dvdalg stream it as a pair list with matrix elements attached.
#include "jci_internal.h"
#include "slater.h"

#include <ctype.h>
#include <math.h>
#include <mutex>
#include <map>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>


static void report(int code, const char *fn)
{
    // source/errorlib.c: "*** Error: <fn>: <code>"
    fprintf(stderr, "*** Error: %s: %d ***\n", fn, code);
    if (jci_last_error()[0]) fprintf(stderr, "    %s\n", jci_last_error());
}

// ---- masks <-> reference strings ----------------------------------------------------------
static uint64_t occ_mask(const occstr &o)
{
    uint64_t m = (uint64_t)o.byte1;
    for (int k = 0; k < o.nvrtx && k < 2; k++)
        if (o.virtx[k] >= 1 && o.virtx[k] <= 64) m |= 1ULL << (o.virtx[k] - 1);
    return m;
}

// byte1 holds orbitals 1..ninto, virtx the (at most two) orbitals above ninto
// (source/binarystr.c:209, str2occstr)
static int make_occ(uint64_t m, int ninto, occstr *o)
{
    memset(o, 0, sizeof(*o));
    uint64_t inner = ninto >= 64 ? ~0ULL : ((1ULL << ninto) - 1ULL);
    o->byte1 = (long long)(m & inner);
    int nv = 0;
    for (uint64_t x = m & ~inner; x; x &= x - 1) {
        if (nv >= 2) return -1;
        o->virtx[nv++] = jci_ctz(x) + 1;
    }
    o->nvrtx = nv;
    return 0;
}

static int max_orbital(const std::vector<uint64_t> &a, const std::vector<uint64_t> &b)
{
    uint64_t all = 0;
    for (uint64_t x : a) all |= x;
    for (uint64_t x : b) all |= x;
    return all ? 64 - __builtin_clzll(all) : 1;
}

static size_t m1_len(int n) { return (size_t)n * (n + 1) / 2; }
static size_t m2_len(int n) { size_t a = m1_len(n); return a * (a + 1) / 2; }

// ---- default device context ---------------------------------------------------------------
static std::mutex g_mu;
static jci_ctx *g_ctx = nullptr;

static int default_ctx(jci_ctx **c)
{
    if (!g_ctx) {
        int rc = jci_ctx_create(0, &g_ctx);
        if (rc) return rc;
    }
    *c = g_ctx;
    return 0;
}

// ---- spaces of the lists citrunc_rtnlist returned ---------------------------------------------
// citrunc_rtnlist registers the space of every list it returns, keyed by the list pointer, so
// that compute_hv and dvdalg on that list run the handle path's sigma (dense-string MFMA for
// full CI, the table kernels otherwise) instead of the explicit pair list.  The key is
// checked on every use (count, and the masks of up to 64 determinants spread over the list
// against the space's own): a list freed and replaced by another at the same address, or a
// caller's own list, falls back to the pair-list path.
struct DropSpace {
    int aelec, belec, orbs, ndocc, nactv, xlvl;
    long long ndets;
    jci_space *sp = nullptr;
    jci_ham *h = nullptr;
    uint64_t key = 0;            // integral hash of h
    const double *km1 = nullptr, *km2 = nullptr;
    std::vector<uint64_t> sa, sb;   // the space's masks at the spot-check determinants
    long long serial = 0;           // registration order
};
static std::map<const void *, DropSpace *> g_spaces;
static long long g_serial = 0;
#define JCI_DROP_SPACES 8           // registered lists kept (the oldest is dropped beyond)

static void drop_space_free(DropSpace *d)
{
    if (!d) return;
    if (d->h) jci_ham_destroy(d->h);
    if (d->sp) jci_space_destroy(d->sp);
    delete d;
}

static uint64_t int_hash(const double *m1, const double *m2, int n);

// the registered space of dlist if it still describes it, with a Hamiltonian for these
// integrals; nullptr otherwise (the caller takes the pair-list path)
static DropSpace *drop_space(struct det *dlist, int ndets, const double *m1, const double *m2)
{
    auto it = g_spaces.find(dlist);
    if (it == g_spaces.end()) return nullptr;
    DropSpace *d = it->second;
    if (d->ndets != ndets) return nullptr;
    jci_ctx *c;
    if (default_ctx(&c)) return nullptr;
    const long long step = ndets > 64 ? ndets / 64 : 1;
    if (!d->sp) {
        if (jci_space_create(c, d->aelec, d->belec, d->orbs, d->ndocc, d->nactv, d->xlvl,
                             JCI_RULE_SERIAL, &d->sp)) { d->sp = nullptr; return nullptr; }
        if (jci_space_ndets(d->sp) != ndets) return nullptr;
        for (long long i = 0; i < ndets; i += step) {
            uint64_t a, b;
            if (jci_space_dets(d->sp, i, 1, &a, &b)) return nullptr;
            d->sa.push_back(a);
            d->sb.push_back(b);
        }
    }
    // spot check: up to 64 determinants spread over the list equal the space's
    size_t t = 0;
    for (long long i = 0; i < ndets && t < d->sa.size(); i += step, t++)
        if (d->sa[t] != occ_mask(dlist[i].astr) || d->sb[t] != occ_mask(dlist[i].bstr)) return nullptr;
    const uint64_t key = int_hash(m1, m2, d->orbs);
    if (!d->h || d->key != key || d->km1 != m1 || d->km2 != m2) {
        if (d->h) { jci_ham_destroy(d->h); d->h = nullptr; }
        if (jci_ham_create(d->sp, m1, m2, JCI_DIRECT_TABLES, &d->h)) { d->h = nullptr; return nullptr; }
        d->key = key;
        d->km1 = m1;
        d->km2 = m2;
    }
    return d;
}

// ---- citrunc_rtnlist (host only) ----------------------------------------------------------
JCI_API struct det *citrunc_rtnlist(int aelec, int belec, int orbs, int nfrzc, int ndocc,
                                    int nactv, int nfrzv, int xlvl, int *astr_len,
                                    int *bstr_len, int *dtrm_len, int *ci_orbs, int *ci_aelec,
                                    int *ci_belec)
{
    *ci_aelec = aelec - nfrzc;
    *ci_belec = belec - nfrzc;
    *ci_orbs = orbs - nfrzc - nfrzv;
    fprintf(stdout, " CI Expansion: %d electrons in %d orbitals\n", *ci_aelec + *ci_belec, *ci_orbs);
    SpaceDesc d;
    if (jci_build_desc(*ci_aelec, *ci_belec, *ci_orbs, ndocc, nactv, xlvl, JCI_RULE_SERIAL, &d)) {
        report(-1, "citrunc_rtnlist");
        return nullptr;
    }
    std::vector<unsigned long long> binom(65 * 65);
    jci_binom_table(binom.data());
    std::vector<uint64_t> masks[2];
    for (int s = 0; s < 2; s++) {
        const SpinDesc &S = d.s[s];
        masks[s].resize(S.gstart[S.ngrp]);
        for (int g = 0; g < S.ngrp; g++)
            for (long long r = 0; r < S.gn[g]; r++) {
                long long nav = S.nsa[g] * S.nsv[g];
                long long rd = r / nav, ra = (r / S.nsv[g]) % S.nsa[g], rv = r % S.nsv[g];
                uint64_t m = jci_lex_unrank(rd, d.ndocc, S.gd[g], binom.data());
                if (S.ga[g]) m |= jci_lex_unrank(ra, d.nactv, S.ga[g], binom.data()) << d.ndocc;
                if (S.gv[g]) m |= jci_lex_unrank(rv, d.nvirt, S.gv[g], binom.data()) << (d.ndocc + d.nactv);
                masks[s][S.gstart[g] + r] = m;
            }
    }
    *astr_len = (int)masks[0].size();
    *bstr_len = (int)masks[1].size();
    fprintf(stdout, " Total number of alpha strings = %d\n", *astr_len);
    fprintf(stdout, " Total number of beta  strings = %d\n", *bstr_len);
    if (d.ndets > 0x7fffffffLL) {
        jci_set_error("citrunc_rtnlist: %lld determinants do not fit an int count", d.ndets);
        report(-1, "citrunc_rtnlist");
        return nullptr;
    }
    *dtrm_len = (int)d.ndets;
    fprintf(stdout, " Number of determinants = %d\n", *dtrm_len);
    struct det *dl = (struct det *)malloc(sizeof(struct det) * (d.ndets > 0 ? d.ndets : 1));
    if (!dl) return nullptr;
    int ninto = ndocc + nactv;
    long long k = 0;
    for (int b = 0; b < d.nblk; b++) {
        int p = d.bp[b], q = d.bq[b];
        int cas = (d.s[0].gv[p] + d.s[1].gv[q]) == 0;
        for (long long i = 0; i < d.s[0].gn[p]; i++)
            for (long long j = 0; j < d.s[1].gn[q]; j++, k++) {
                if (make_occ(masks[0][d.s[0].gstart[p] + i], ninto, &dl[k].astr) ||
                    make_occ(masks[1][d.s[1].gstart[q] + j], ninto, &dl[k].bstr)) {
                    free(dl);
                    jci_set_error("citrunc_rtnlist: more than two virtual orbitals in a string");
                    report(-1, "citrunc_rtnlist");
                    return nullptr;
                }
                dl[k].cas = cas;
            }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    DropSpace *ds = new DropSpace();
    ds->aelec = *ci_aelec; ds->belec = *ci_belec; ds->orbs = *ci_orbs;
    ds->ndocc = ndocc; ds->nactv = nactv; ds->xlvl = xlvl;
    ds->ndets = d.ndets;
    ds->serial = ++g_serial;
    auto it = g_spaces.find(dl);
    if (it != g_spaces.end()) { drop_space_free(it->second); g_spaces.erase(it); }
    if (g_spaces.size() >= JCI_DROP_SPACES) {
        auto old = g_spaces.begin();
        for (auto jt = g_spaces.begin(); jt != g_spaces.end(); ++jt)
            if (jt->second->serial < old->second->serial) old = jt;
        drop_space_free(old->second);
        g_spaces.erase(old);
    }
    g_spaces[dl] = ds;
    return dl;
}

// ---- explicit-list operator -------------------------------------------------------------------
struct ListOp {
    jci_ctx *c;
    long long n;
    int norbs;
    std::vector<uint64_t> ha, hb;
    uint64_t *d_ma = nullptr, *d_mb = nullptr;
    std::vector<long long> up_ptr;   // upper pattern incl. diagonal (host)
    std::vector<int> up_col;
    // matrix elements for one set of integrals
    const double *key_m1 = nullptr, *key_m2 = nullptr;
    uint64_t key_sum = 0;
    double *d_m1 = nullptr, *d_m2 = nullptr, *d_hd = nullptr, *d_val = nullptr;
    long long *d_rowptr = nullptr;
    int *d_col = nullptr;
    long long nnz = 0;
    double *d_x = nullptr, *d_y = nullptr;
};

static std::map<const void *, ListOp *> g_ops;

static void free_op(ListOp *op)
{
    if (!op) return;
    hipSetDevice(op->c->device);
    hipStreamSynchronize(op->c->stream);
    hipFree(op->d_ma); hipFree(op->d_mb); hipFree(op->d_m1); hipFree(op->d_m2);
    hipFree(op->d_hd); hipFree(op->d_val); hipFree(op->d_rowptr); hipFree(op->d_col);
    hipFree(op->d_x); hipFree(op->d_y);
    delete op;
}

static int op_masks(ListOp *op, struct det *dlist, long long n)
{
    op->n = n;
    op->ha.resize(n);
    op->hb.resize(n);
    for (long long i = 0; i < n; i++) {
        op->ha[i] = occ_mask(dlist[i].astr);
        op->hb[i] = occ_mask(dlist[i].bstr);
    }
    op->norbs = max_orbital(op->ha, op->hb);
    JCI_CHECK(hipMalloc(&op->d_ma, sizeof(uint64_t) * n));
    JCI_CHECK(hipMalloc(&op->d_mb, sizeof(uint64_t) * n));
    JCI_CHECK(hipMemcpyAsync(op->d_ma, op->ha.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice,
                             op->c->stream));
    JCI_CHECK(hipMemcpyAsync(op->d_mb, op->hb.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice,
                             op->c->stream));
    return 0;
}

// the O(N^2) screen on the device: upper pattern (j >= i) in op->up_ptr/up_col
static int op_screen(ListOp *op)
{
    long long n = op->n;
    long long *d_cnt, *d_ptr;
    int *d_col;
    JCI_CHECK(hipMalloc(&d_cnt, sizeof(long long) * n));
    JCI_CHECK(hipMalloc(&d_ptr, sizeof(long long) * (n + 1)));
    int rc = jk_list_screen_count(n, op->d_ma, op->d_mb, d_cnt, op->c->stream);
    if (rc) return rc;
    rc = jk_scan(d_cnt, d_ptr, n, op->c->stream);
    if (rc) return rc;
    op->up_ptr.resize(n + 1);
    JCI_CHECK(hipMemcpyAsync(op->up_ptr.data(), d_ptr, sizeof(long long) * (n + 1),
                             hipMemcpyDeviceToHost, op->c->stream));
    JCI_CHECK(hipStreamSynchronize(op->c->stream));
    long long nz = op->up_ptr[n];
    JCI_CHECK(hipMalloc(&d_col, sizeof(int) * (nz > 0 ? nz : 1)));
    rc = jk_list_screen_fill(n, op->d_ma, op->d_mb, d_ptr, d_col, op->c->stream);
    if (rc) return rc;
    op->up_col.resize(nz);
    JCI_CHECK(hipMemcpyAsync(op->up_col.data(), d_col, sizeof(int) * nz, hipMemcpyDeviceToHost,
                             op->c->stream));
    JCI_CHECK(hipStreamSynchronize(op->c->stream));
    hipFree(d_cnt); hipFree(d_ptr); hipFree(d_col);
    return 0;
}

// upper pattern from a caller-built hmap (reference section lists)
static void op_pattern_from_hmap(ListOp *op, const struct rowmap *hmap)
{
    long long n = op->n;
    op->up_ptr.assign(n + 1, 0);
    op->up_col.clear();
    for (long long i = 0; i < n; i++) {
        op->up_ptr[i] = (long long)op->up_col.size();
        for (const struct section *s = hmap[i].sec; s; s = s->next)
            for (long long j = s->first + i; j <= s->last + i && j < n; j++)
                op->up_col.push_back((int)j);
    }
    op->up_ptr[n] = (long long)op->up_col.size();
}

// 64-bit hash of every word of the integral arrays (bit patterns, so any change to any
// integral changes the key): the cached matrix elements are reused only for the same values
static uint64_t int_hash(const double *m1, const double *m2, int n)
{
    uint64_t h = 0x9e3779b97f4a7c15ULL;
    auto mix = [&](const double *a, size_t len) {
        for (size_t i = 0; i < len; i++) {
            uint64_t w;
            memcpy(&w, a + i, sizeof(w));
            h ^= w + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
            h *= 0xff51afd7ed558ccdULL;
            h ^= h >> 33;
        }
    };
    mix(m1, m1_len(n));
    mix(m2, m2_len(n));
    return h;
}

// matrix elements of the full pattern (both triangles, diagonal separate) for m1/m2
static int op_values(ListOp *op, const double *m1, const double *m2)
{
    int n_orb = op->norbs;
    const uint64_t sum = int_hash(m1, m2, n_orb);
    if (op->d_val && op->key_m1 == m1 && op->key_m2 == m2 && op->key_sum == sum) return 0;
    jci_ctx *c = op->c;
    long long n = op->n;
    hipFree(op->d_m1); hipFree(op->d_m2);
    JCI_CHECK(hipMalloc(&op->d_m1, sizeof(double) * m1_len(n_orb)));
    JCI_CHECK(hipMalloc(&op->d_m2, sizeof(double) * m2_len(n_orb)));
    JCI_CHECK(hipMemcpyAsync(op->d_m1, m1, sizeof(double) * m1_len(n_orb), hipMemcpyHostToDevice, c->stream));
    JCI_CHECK(hipMemcpyAsync(op->d_m2, m2, sizeof(double) * m2_len(n_orb), hipMemcpyHostToDevice, c->stream));
    if (!op->d_hd) JCI_CHECK(hipMalloc(&op->d_hd, sizeof(double) * n));
    int rc = jk_list_hdiag(n, op->d_ma, op->d_mb, n_orb, op->d_m1, op->d_m2, op->d_hd, c->stream);
    if (rc) return rc;
    if (!op->d_col) {
        // symmetric completion of the upper pattern, rows sorted ascending
        std::vector<long long> cnt(n + 1, 0);
        for (long long i = 0; i < n; i++)
            for (long long t = op->up_ptr[i]; t < op->up_ptr[i + 1]; t++) {
                long long j = op->up_col[t];
                if (j == i) continue;
                cnt[i]++;
                cnt[j]++;
            }
        std::vector<long long> ptr(n + 1, 0);
        for (long long i = 0; i < n; i++) ptr[i + 1] = ptr[i] + cnt[i];
        op->nnz = ptr[n];
        std::vector<int> col(op->nnz > 0 ? op->nnz : 1);
        std::vector<long long> pos(ptr.begin(), ptr.end() - 1);
        for (long long i = 0; i < n; i++)
            for (long long t = op->up_ptr[i]; t < op->up_ptr[i + 1]; t++) {
                long long j = op->up_col[t];
                if (j == i) continue;
                col[pos[i]++] = (int)j;
                col[pos[j]++] = (int)i;
            }
        JCI_CHECK(hipMalloc(&op->d_rowptr, sizeof(long long) * (n + 1)));
        JCI_CHECK(hipMalloc(&op->d_col, sizeof(int) * col.size()));
        JCI_CHECK(hipMalloc(&op->d_val, sizeof(double) * col.size()));
        JCI_CHECK(hipMemcpyAsync(op->d_rowptr, ptr.data(), sizeof(long long) * (n + 1),
                                 hipMemcpyHostToDevice, c->stream));
        JCI_CHECK(hipMemcpyAsync(op->d_col, col.data(), sizeof(int) * col.size(),
                                 hipMemcpyHostToDevice, c->stream));
        JCI_CHECK(hipStreamSynchronize(c->stream));
    }
    rc = jk_list_values(op->nnz, n, op->d_ma, op->d_mb, n_orb, op->d_m1, op->d_m2, op->d_rowptr,
                        op->d_col, op->d_val, c->stream);
    if (rc) return rc;
    JCI_CHECK(hipStreamSynchronize(c->stream));
    op->key_m1 = m1;
    op->key_m2 = m2;
    op->key_sum = sum;
    return 0;
}

// operator for hmap: registered by generate_cimap, or built from a foreign hmap
static int get_op(struct det *dlist, int ndets, struct rowmap *hmap, ListOp **out)
{
    auto it = g_ops.find(hmap);
    if (it != g_ops.end() && it->second->n == ndets) { *out = it->second; return 0; }
    jci_ctx *c;
    int rc = default_ctx(&c);
    if (rc) return rc;
    ListOp *op = new ListOp();
    op->c = c;
    if ((rc = op_masks(op, dlist, ndets))) { free_op(op); return rc; }
    if (hmap) op_pattern_from_hmap(op, hmap);
    else if ((rc = op_screen(op))) { free_op(op); return rc; }
    if (it != g_ops.end()) free_op(it->second);
    g_ops[hmap] = op;
    *out = op;
    return 0;
}

JCI_API int generate_cimap(struct det *dlist, int ndets, int nactv, struct rowmap *hmap)
{
    (void)nactv;
    std::lock_guard<std::mutex> lk(g_mu);
    jci_ctx *c;
    int rc = default_ctx(&c);
    if (rc) { report(rc, "generate_cimap"); return rc; }
    JCI_CHECK(hipSetDevice(c->device));
    ListOp *op = new ListOp();
    op->c = c;
    if ((rc = op_masks(op, dlist, ndets)) || (rc = op_screen(op))) {
        free_op(op);
        report(rc, "generate_cimap");
        return rc;
    }
    // reference section lists: maximal runs of consecutive connected j >= i, stored as
    // offsets from i (source/cimapping.c:23-65)
    for (long long i = 0; i < ndets; i++) {
        hmap[i].nsec = 0;
        hmap[i].sec = nullptr;
        struct section *prev = nullptr;
        long long t = op->up_ptr[i], e = op->up_ptr[i + 1];
        while (t < e) {
            long long j0 = op->up_col[t], j1 = j0;
            while (t + 1 < e && op->up_col[t + 1] == j1 + 1) { t++; j1++; }
            t++;
            struct section *s = (struct section *)malloc(sizeof(struct section));
            s->first = (int)(j0 - i);
            s->last = (int)(j1 - i);
            s->next = nullptr;
            s->prev = prev;
            if (prev) prev->next = s; else hmap[i].sec = s;
            prev = s;
            hmap[i].nsec++;
        }
    }
    auto it = g_ops.find(hmap);
    if (it != g_ops.end()) free_op(it->second);
    g_ops[hmap] = op;
    return 0;
}

JCI_API int deallocate_cimap(struct rowmap *hmap, int ndets)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ops.find(hmap);
    if (it != g_ops.end()) { free_op(it->second); g_ops.erase(it); }
    int i;
    for (i = 0; i < ndets; i++) {
        struct section *s = hmap[i].sec;
        while (s) { struct section *nx = s->next; free(s); s = nx; }
    }
    free(hmap);
    return i;
}

JCI_API double hmatels(struct det deti, struct det detj, double *moints1, double *moints2,
                       int aelec, int belec, int nactv)
{
    (void)aelec; (void)belec; (void)nactv;
    int dg, ph;
    return jci_hij(occ_mask(deti.astr), occ_mask(deti.bstr), occ_mask(detj.astr),
                   occ_mask(detj.bstr), moints1, moints2, &dg, &ph);
}

JCI_API void compute_hdgls(struct det *dlist, int ndets, double *moints1, double *moints2,
                           int aelec, int belec, double *hdgls, int ninto)
{
    (void)aelec; (void)belec; (void)ninto;
    std::lock_guard<std::mutex> lk(g_mu);
    jci_ctx *c;
    int rc = default_ctx(&c);
    if (rc) { report(rc, "compute_hdgls"); return; }
    ListOp tmp;
    tmp.c = c;
    if ((rc = op_masks(&tmp, dlist, ndets)) == 0) {
        int n_orb = tmp.norbs;
        double *d_m1 = nullptr, *d_m2 = nullptr, *d_hd = nullptr;
        if (hipMalloc(&d_m1, sizeof(double) * m1_len(n_orb)) != hipSuccess ||
            hipMalloc(&d_m2, sizeof(double) * m2_len(n_orb)) != hipSuccess ||
            hipMalloc(&d_hd, sizeof(double) * ndets) != hipSuccess) rc = -3;
        if (!rc) {
            hipMemcpyAsync(d_m1, moints1, sizeof(double) * m1_len(n_orb), hipMemcpyHostToDevice, c->stream);
            hipMemcpyAsync(d_m2, moints2, sizeof(double) * m2_len(n_orb), hipMemcpyHostToDevice, c->stream);
            rc = jk_list_hdiag(ndets, tmp.d_ma, tmp.d_mb, n_orb, d_m1, d_m2, d_hd, c->stream);
            if (!rc && (hipMemcpyAsync(hdgls, d_hd, sizeof(double) * ndets, hipMemcpyDeviceToHost,
                                       c->stream) != hipSuccess ||
                        hipStreamSynchronize(c->stream) != hipSuccess)) rc = -2;
        }
        hipFree(d_m1); hipFree(d_m2); hipFree(d_hd);
    }
    hipFree(tmp.d_ma); hipFree(tmp.d_mb);
    tmp.d_ma = tmp.d_mb = nullptr;
    if (rc) report(rc, "compute_hdgls");
}

JCI_API void compute_hv(struct det *dlist, int ndets, double *moints1, double *moints2,
                        int aelec, int belec, double *v, double *c, int ninto,
                        struct rowmap *hmap)
{
    (void)aelec; (void)belec; (void)ninto;
    std::lock_guard<std::mutex> lk(g_mu);
    if (DropSpace *ds = drop_space(dlist, ndets, moints1, moints2)) {
        // the space's own sigma on the device; the reference accumulates: c += H v
        jci_ctx *cx;
        default_ctx(&cx);
        double *dx = nullptr, *dy = nullptr;
        std::vector<double> y(ndets);
        int rc2 = 0;
        if (hipMalloc(&dx, sizeof(double) * ndets) != hipSuccess ||
            hipMalloc(&dy, sizeof(double) * ndets) != hipSuccess) rc2 = -3;
        if (!rc2 && hipMemcpyAsync(dx, v, sizeof(double) * ndets, hipMemcpyHostToDevice,
                                   cx->stream) != hipSuccess) rc2 = -2;
        if (!rc2) rc2 = jci_sigma_dev(ds->h, JCI_SIGMA_AUTO, 1, dx, dy);
        if (!rc2 && (hipMemcpyAsync(y.data(), dy, sizeof(double) * ndets, hipMemcpyDeviceToHost,
                                    cx->stream) != hipSuccess ||
                     hipStreamSynchronize(cx->stream) != hipSuccess)) rc2 = -2;
        hipFree(dx);
        hipFree(dy);
        if (rc2) { report(rc2, "compute_hv"); return; }
        for (int i = 0; i < ndets; i++) c[i] += y[i];
        return;
    }
    ListOp *op;
    int rc = get_op(dlist, ndets, hmap, &op);
    if (!rc) rc = op_values(op, moints1, moints2);
    if (rc) { report(rc, "compute_hv"); return; }
    jci_ctx *cx = op->c;
    long long n = op->n;
    if (!op->d_x) {
        if (hipMalloc(&op->d_x, sizeof(double) * n) != hipSuccess ||
            hipMalloc(&op->d_y, sizeof(double) * n) != hipSuccess) { report(-3, "compute_hv"); return; }
    }
    std::vector<double> y(n);
    if (hipMemcpyAsync(op->d_x, v, sizeof(double) * n, hipMemcpyHostToDevice, cx->stream) != hipSuccess) {
        report(-2, "compute_hv");
        return;
    }
    rc = jk_spmv(64, 1, n, op->d_rowptr, op->d_col, op->d_val, op->d_hd, op->d_x, op->d_y, cx->stream);
    if (!rc && (hipMemcpyAsync(y.data(), op->d_y, sizeof(double) * n, hipMemcpyDeviceToHost,
                               cx->stream) != hipSuccess ||
                hipStreamSynchronize(cx->stream) != hipSuccess)) rc = -2;
    if (rc) { report(rc, "compute_hv"); return; }
    // the reference accumulates: c += H v
    for (long long i = 0; i < n; i++) c[i] += y[i];
}

// ---- dvdalg -----------------------------------------------------------------------------------
static int list_sigma(void *self, int method, int k, const double *x, double *y)
{
    (void)method;
    ListOp *op = (ListOp *)self;
    return jk_spmv(64, k, op->n, op->d_rowptr, op->d_col, op->d_val, op->d_hd, x, y, op->c->stream);
}

static int list_masks(void *self, long long n, const long long *d_idx, uint64_t *oa, uint64_t *ob)
{
    ListOp *op = (ListOp *)self;
    std::vector<long long> idx(n);
    JCI_CHECK(hipMemcpyAsync(idx.data(), d_idx, sizeof(long long) * n, hipMemcpyDeviceToHost,
                             op->c->stream));
    JCI_CHECK(hipStreamSynchronize(op->c->stream));
    std::vector<uint64_t> a(n), b(n);
    for (long long t = 0; t < n; t++) { a[t] = op->ha[idx[t]]; b[t] = op->hb[idx[t]]; }
    JCI_CHECK(hipMemcpyAsync(oa, a.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, op->c->stream));
    JCI_CHECK(hipMemcpyAsync(ob, b.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, op->c->stream));
    JCI_CHECK(hipStreamSynchronize(op->c->stream));
    return 0;
}

static int list_resolve(void *, int, int) { return JCI_SIGMA_PAIRS64; }

JCI_API int dvdalg(struct det *dlist, int ndets, double *moints1, double *moints2, int aelec,
                   int belec, double *hdgls, int ninto, double totfrze, int maxiter, int krymin,
                   int krymax, int nroots, double restol, struct rowmap *hmap, double **civec,
                   double *cival, int predr, int plvl, int norbs, int refdim)
{
    (void)ninto;
    std::lock_guard<std::mutex> lk(g_mu);
    if (predr != 1 && predr != 2) {
        fprintf(stderr, "No initial guess routine selected!\n");
        report(-1, "dvdalg");
        return -1;
    }
    DropSpace *ds = drop_space(dlist, ndets, moints1, moints2);
    ListOp *op = nullptr;
    int rc = 0;
    if (!ds) {
        rc = get_op(dlist, ndets, hmap, &op);
        if (!rc) rc = op_values(op, moints1, moints2);
        if (rc) { report(rc, "dvdalg"); return rc; }
    }
    jci_ctx *octx;
    if ((rc = default_ctx(&octx))) { report(rc, "dvdalg"); return rc; }
    jci_solve_params p;
    jci_solve_defaults(&p);
    p.maxiter = maxiter;
    p.krymin = krymin;
    p.krymax = krymax;
    p.nroots = nroots;
    p.restol = restol;
    p.refdim = refdim;
    p.totfrze = totfrze;
    p.verbose = plvl > 1;
    p.precision = JCI_PREC_FP64;
    std::vector<double> guess;
    if (predr == 2) {
        // source/initguess_roldv.c:20: krymin vectors from civector.dat
        std::vector<double> store((size_t)krymin * ndets), ev(krymin);
        std::vector<double *> rows(krymin);
        for (int r = 0; r < krymin; r++) rows[r] = store.data() + (size_t)r * ndets;
        rc = read_civectors(aelec, belec, norbs, rows.data(), ndets, krymin, ev.data());
        if (rc) {
            // source/initguess_roldv.c:30-41: 1 (no file) and 2 (header) pass through, any
            // other reader code becomes 100
            if (rc != 1 && rc != 2) rc = 100;
            report(rc, "dvdalg");
            return rc;
        }
        guess.swap(store);
        p.guess = JCI_GUESS_USER;
    } else {
        p.guess = JCI_GUESS_BLOCK;
    }
    // preconditioner diagonal: the caller's hdgls
    double *d_pre = nullptr;
    if (hdgls) {
        JCI_CHECK(hipMalloc(&d_pre, sizeof(double) * ndets));
        JCI_CHECK(hipMemcpyAsync(d_pre, hdgls, sizeof(double) * ndets, hipMemcpyHostToDevice,
                                 octx->stream));
    }
    SolveOp so{};
    if (ds) {
        // the space's own sigma (dense-string MFMA for full CI, table kernels otherwise)
        jci_ham_solve_op(ds->h, d_pre, &so);
    } else {
        so.self = op;
        so.ctx = op->c;
        so.N = op->n;
        so.d_hd = d_pre ? d_pre : op->d_hd;
        so.d_m1 = op->d_m1;
        so.d_m2 = op->d_m2;
        so.sigma = list_sigma;
        so.masks = list_masks;
        so.resolve = list_resolve;
    }
    std::vector<double> ev(nroots), vec((size_t)nroots * ndets);
    jci_solve_stats st;
    rc = jci_solve_core(&so, &p, guess.empty() ? nullptr : guess.data(), ev.data(), vec.data(), &st, 0);
    hipFree(d_pre);
    if (rc < 0) { report(rc, "dvdalg"); return rc; }
    if (rc == 0) fprintf(stdout, "\n ** CI CONVERGED! **\n");
    fprintf(stdout, " Davidson algorithm finished.\n");
    for (int r = 0; r < nroots; r++) {
        cival[r] = ev[r];
        memcpy(civec[r], vec.data() + (size_t)r * ndets, sizeof(double) * ndets);
    }
    // like the reference, running out of iterations is not an error code
    return 0;
}

// ---- the MPI driver's binary vector files and its Davidson step --------------------------------
// Layout of mpi_source/mpi_utilities.c:353-415 (writer) and :426-531 (reader): int32 dim,
// int32 len, then dim vectors of len doubles, native byte order, no record markers.
JCI_API int jci_write_vectors_ufmt(const char *fname, int len, int dim, double *const *vecs)
{
    if (len < 0 || dim < 0 || (dim > 0 && !vecs)) { jci_set_error("write_vectors_ufmt: bad arguments"); return -1; }
    FILE *f = fopen(fname, "wb");
    if (!f) { jci_set_error("write_vectors_ufmt: cannot open %s", fname); return 2; }
    const int32_t hdr[2] = {dim, len};
    int rc = fwrite(hdr, sizeof(int32_t), 2, f) == 2 ? 0 : 6;
    for (int i = 0; i < dim && rc == 0; i++)
        if (fwrite(vecs[i], sizeof(double), (size_t)len, f) != (size_t)len) rc = 6;
    if (fclose(f) != 0 && rc == 0) rc = 6;
    if (rc) jci_set_error("write_vectors_ufmt: write error on %s", fname);
    return rc;
}

JCI_API int jci_read_vectors_ufmt(const char *fname, int len, int dim, double **vecs)
{
    if (len < 0 || dim < 0 || (dim > 0 && !vecs)) { jci_set_error("read_vectors_ufmt: bad arguments"); return -1; }
    FILE *f = fopen(fname, "rb");
    if (!f) { jci_set_error("read_vectors_ufmt: cannot open %s", fname); return 2; }
    int32_t fdim = 0, flen = 0;
    int rc = 0;
    if (fread(&fdim, sizeof(int32_t), 1, f) != 1) rc = 3;
    else if (fread(&flen, sizeof(int32_t), 1, f) != 1) rc = 4;
    else if (fdim != dim || flen != len) rc = 5;
    for (int i = 0; i < dim && rc == 0; i++)
        if (fread(vecs[i], sizeof(double), (size_t)len, f) != (size_t)len) rc = 6;
    fclose(f);
    if (rc == 5)
        jci_set_error("read_vectors_ufmt: %s holds %d x %d, expected %d x %d", fname, fdim, flen, dim, len);
    else if (rc)
        jci_set_error("read_vectors_ufmt: %s is short (code %d)", fname, rc);
    return rc;
}

static std::string in_dir(const char *dir, const char *name)
{
    if (!dir || !*dir) return name;
    std::string s(dir);
    if (s.back() != '/') s += '/';
    return s + name;
}

// mpi_source/pdavidson.c:39-356 on a handle: guess by prediagr (:183, :385-449), the solve,
// then civec and, unconverged, ci.restart and hvvec (:343-347)
JCI_API int jci_pdavidson(jci_ham *h, const jci_solve_params *p, int prediagr, const char *dir,
                          double *evals, jci_solve_stats *stats)
{
    if (!h || !p) { jci_set_error("pdavidson: null argument"); return -1; }
    const long long N = jci_ham_ndets(h);
    if (N > INT32_MAX) { jci_set_error("pdavidson: %lld determinants exceed the file format", N); return -1; }
    if (p->nroots < 1 || p->krymin < p->nroots) { jci_set_error("pdavidson: need 1 <= nroots <= krymin"); return -1; }
    jci_solve_params q = *p;
    std::vector<double> guess;
    if (prediagr == 1) {
        q.guess = JCI_GUESS_BLOCK;
    } else if (prediagr == 2) {
        if (p->krymin > p->refdim) { jci_set_error("pdavidson: unit-vector guess needs krymin <= refdim"); return -1; }
        q.guess = JCI_GUESS_UNIT;
    } else if (prediagr == 3) {
        guess.resize((size_t)p->krymin * N);
        std::vector<double *> rows(p->krymin);
        for (int r = 0; r < p->krymin; r++) rows[r] = guess.data() + (size_t)r * N;
        int rc = jci_read_vectors_ufmt(in_dir(dir, "ci.restart").c_str(), (int)N, p->krymin, rows.data());
        if (rc) return rc;
        q.guess = JCI_GUESS_USER;
    } else {
        jci_set_error("pdavidson: prediagr must be 1, 2 or 3");
        return -1;
    }
    const int nout = std::max(p->nroots, p->krymin);
    std::vector<double> ev(p->nroots), vec((size_t)nout * N);
    jci_solve_stats st;
    int rc = jci_solve_nout(h, &q, guess.empty() ? nullptr : guess.data(), ev.data(), vec.data(), &st, nout);
    if (stats) *stats = st;
    if (rc < 0) return rc;
    if (evals) for (int r = 0; r < p->nroots; r++) evals[r] = ev[r];
    std::vector<double *> rows(nout);
    for (int r = 0; r < nout; r++) rows[r] = vec.data() + (size_t)r * N;
    int wrc = jci_write_vectors_ufmt(in_dir(dir, "civec").c_str(), (int)N, p->nroots, rows.data());
    if (!wrc && rc == 1) {
        wrc = jci_write_vectors_ufmt(in_dir(dir, "ci.restart").c_str(), (int)N, p->krymin, rows.data());
        // the reference clears its sigma block before the loop ends, so hvvec holds zeros
        std::vector<double> zero(N, 0.0);
        std::vector<double *> zrows(p->nroots, zero.data());
        if (!wrc) wrc = jci_write_vectors_ufmt(in_dir(dir, "hvvec").c_str(), (int)N, p->nroots, zrows.data());
    }
    if (wrc) return wrc;
    return rc;
}

// ---- output files -------------------------------------------------------------------------------
JCI_API int print_civectors(int aelec, int belec, int orbs, double **civec, int ndets, int roots,
                            double *cival)
{
    FILE *f = fopen("civector.dat", "w");
    if (!f) { report(1, "print_civectors"); return 1; }
    fprintf(f, "%d %d %d\n", aelec + belec, orbs, ndets);
    fprintf(f, "%d\n\n", roots);
    for (int i = 0; i < roots; i++) {
        fprintf(f, "%15.8lf\n\n", cival[i]);
        for (int j = 0; j < ndets; j++) fprintf(f, "%15.12lf\n", civec[i][j]);
        fprintf(f, "\n");
    }
    fclose(f);
    return 0;
}

JCI_API int read_civectors(int aelec, int belec, int orbs, double **civec, int ndets, int roots,
                           double *cival)
{
    FILE *f = fopen("civector.dat", "r");
    if (!f) { report(1, "read_civectors"); return 1; }
    int te = 0, fo = 0, fn = 0, fr = 0;
    if (fscanf(f, "%d %d %d", &te, &fo, &fn) != 3 || te != aelec + belec || fo != orbs ||
        fn != ndets) {
        fclose(f);
        fprintf(stderr, "*** Error: read_civectors: incorrect parameters ***\n");
        report(2, "read_civectors");
        return 2;
    }
    if (fscanf(f, "%d", &fr) != 1 || fr < roots) {
        fclose(f);
        fprintf(stderr, "*** Error: read_civectors: File contains less roots than desired ***\n");
        report(10 + fr, "read_civectors");
        return 10 + fr;
    }
    for (int i = 0; i < roots; i++) {
        if (fscanf(f, "%lf", &cival[i]) != 1) { fclose(f); report(3, "read_civectors"); return 3; }
        for (int j = 0; j < ndets; j++)
            if (fscanf(f, "%lf", &civec[i][j]) != 1) { fclose(f); report(3, "read_civectors"); return 3; }
    }
    fclose(f);
    return 0;
}

// ---- namelists -----------------------------------------------------------------------------------
// Minimal Fortran namelist reader for "&group key = value, ... /" with integer and real
// values (d/D exponents accepted).  Keys are case-insensitive.
static int read_group(const char *path, const char *group, std::map<std::string, std::string> &kv)
{
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    fclose(f);
    std::string low = text;
    for (auto &ch : low) ch = (char)tolower((unsigned char)ch);
    std::string tag = std::string("&") + group;
    size_t pos = 0;
    for (;;) {
        pos = low.find(tag, pos);
        if (pos == std::string::npos) return 2;
        size_t after = pos + tag.size();
        if (after >= low.size() || !(isalnum((unsigned char)low[after]) || low[after] == '_')) break;
        pos = after;
    }
    size_t end = low.find('/', pos);
    if (end == std::string::npos) end = low.size();
    std::string body = low.substr(pos + tag.size(), end - pos - tag.size());
    // strip comments (! to end of line)
    std::string clean;
    bool comment = false;
    for (char ch : body) {
        if (ch == '!') comment = true;
        if (ch == '\n') comment = false;
        if (!comment) clean.push_back(ch == ',' || ch == '\n' || ch == '\r' || ch == '\t' ? ' ' : ch);
    }
    size_t i = 0;
    while (i < clean.size()) {
        while (i < clean.size() && clean[i] == ' ') i++;
        size_t k0 = i;
        while (i < clean.size() && clean[i] != '=' && clean[i] != ' ') i++;
        std::string key = clean.substr(k0, i - k0);
        while (i < clean.size() && clean[i] == ' ') i++;
        if (i >= clean.size() || clean[i] != '=') { if (key.empty()) break; continue; }
        i++;
        while (i < clean.size() && clean[i] == ' ') i++;
        size_t v0 = i;
        while (i < clean.size() && clean[i] != ' ') i++;
        std::string val = clean.substr(v0, i - v0);
        for (auto &ch : val) if (ch == 'd') ch = 'e';
        if (!key.empty()) kv[key] = val;
    }
    return 0;
}

static void get_int(const std::map<std::string, std::string> &kv, const char *k, int *out)
{
    auto it = kv.find(k);
    if (it != kv.end()) *out = (int)strtol(it->second.c_str(), nullptr, 10);
}

static void get_dbl(const std::map<std::string, std::string> &kv, const char *k, double *out)
{
    auto it = kv.find(k);
    if (it != kv.end()) *out = strtod(it->second.c_str(), nullptr);
}

JCI_API void readgeninput(int *elec, int *orbs, int *nfrozen, int *ndocc, int *nactive,
                          int *xlevel, int *nfrzvirt, int *printlvl, int *printwvf, int *err)
{
    // defaults of source/readnamelist.f90:50-57
    *elec = 0; *orbs = 0; *nfrozen = 0; *ndocc = 0; *nactive = 0; *xlevel = 0; *nfrzvirt = 0;
    *printlvl = 0; *printwvf = 0;
    std::map<std::string, std::string> kv;
    *err = read_group("jayci.in", "general", kv);
    if (*err) return;
    get_int(kv, "electrons", elec);
    get_int(kv, "orbitals", orbs);
    get_int(kv, "nfrozen", nfrozen);
    get_int(kv, "ndocc", ndocc);
    get_int(kv, "nactive", nactive);
    get_int(kv, "xlevel", xlevel);
    get_int(kv, "nfrzvirt", nfrzvirt);
    get_int(kv, "printlvl", printlvl);
    get_int(kv, "printwvf", printwvf);
}

JCI_API void readdaiinput(int *maxiter, int *krymin, int *krymax, int *nroots, int *prediagr,
                          int *refdim, double *restol, int *err)
{
    // defaults of source/readnamelist.f90:79-85; restol kept at full precision
    *maxiter = 20; *krymin = 3; *krymax = 8; *nroots = 1; *prediagr = 1; *restol = 1.0e-5;
    *refdim = 3;
    std::map<std::string, std::string> kv;
    *err = read_group("jayci.in", "dalginfo", kv);
    if (*err) return;
    get_int(kv, "maxiter", maxiter);
    get_int(kv, "krymin", krymin);
    get_int(kv, "krymax", krymax);
    get_int(kv, "nroots", nroots);
    get_int(kv, "prediagr", prediagr);
    get_int(kv, "refdim", refdim);
    get_dbl(kv, "restol", restol);
}

JCI_API void write_wavefunction(struct det *detlist, int ndets, int nroots, double **civec,
                                double *cival, int orbs, int ninto, int elec)
{
    int electrons, orbitals, nfrzncore, ndocc, nactive, xlevel, nfrznvirt, printlvl, printwvf, err;
    readgeninput(&electrons, &orbitals, &nfrzncore, &ndocc, &nactive, &xlevel, &nfrznvirt,
                 &printlvl, &printwvf, &err);
    if (err) { fprintf(stderr, "*** Error: write_wavefunction: Could not input file. ***\n"); return; }
    FILE *f = fopen("ciwvfcn.dat", "w");
    if (!f) { fprintf(stderr, "*** Error: write_wavefunction: Could not open ciwvfcn.dat! ***\n"); return; }
    fprintf(f, "%d %d\n", electrons, orbitals);
    fprintf(f, "%d %d %d %d\n", nfrzncore, ndocc, nactive, nfrznvirt);
    fprintf(f, "%d\n", xlevel);
    fprintf(f, "%d %d %d %d\n\n", elec, orbs, ninto, ndets);
    for (int i = 0; i < ndets; i++) {
        const struct det &d = detlist[i];
        fprintf(f, "%lld %lld %d %d %d  %lld %lld %d %d %d  %d\n", d.astr.byte1, d.astr.byte2,
                d.astr.virtx[0], d.astr.virtx[1], d.astr.nvrtx, d.bstr.byte1, d.bstr.byte2,
                d.bstr.virtx[0], d.bstr.virtx[1], d.bstr.nvrtx, d.cas);
    }
    fprintf(f, "\n\n");
    fprintf(f, "%d\n\n", nroots);
    for (int i = 0; i < nroots; i++) {
        fprintf(f, "%15.8lf\n\n", cival[i]);
        for (int j = 0; j < ndets; j++) fprintf(f, "%15.12lf\n", civec[i][j]);
        fprintf(f, "\n");
    }
    fclose(f);
}
